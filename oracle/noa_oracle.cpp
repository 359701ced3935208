/*
 * TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * CPU restatement of noa's central-slice operators (thomasfrosio/noa v0.10.0). Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library: as the parity checker of the
 * HIP path and as the timed CPU baseline. Nothing in noa_amd/ links or calls it.
 *
 * It follows, expression by expression and in the same evaluation order, the reference's CPU path:
 *   - FourierInsertRasterize   src/noa/xform/FourierProject.hpp:161-390
 *   - FourierInsertInterpolate src/noa/xform/FourierProject.hpp:392-539
 *   - FourierExtract (3D, 4D)  src/noa/xform/FourierProject.hpp:541-725
 *   - FourierInsertExtract     src/noa/xform/FourierProject.hpp:727-1006 (launcher :1489-1598)
 *   - CTFAnisotropic, CTF op   src/noa/signal/core/CTF.hpp:12-25, :291-575; src/noa/signal/CTF.hpp:14-125
 *   - windowed sinc helpers    src/noa/xform/FourierProject.hpp:97-159
 *   - slice<->grid transforms  src/noa/xform/FourierProject.hpp:38-95
 *   - interpolate_spectrum     src/noa/xform/core/Interpolation.hpp:416-526 (+ weights :64-146): NEAREST, LINEAR,
 *                              CUBIC, CUBIC_BSPLINE, LANCZOS4/6/8
 *   - index/frequency math     src/noa/fft/core/Frequency.hpp:80-145, :250-288
 *   - CPU iwise execution      src/noa/runtime/cpu/Iwise.hpp:7-117 (+ atomics cpu/ComputeHandle.hpp:39-55)
 *   - parameter defaults       src/noa/xform/FourierProject.hpp:1260-1487 (launchers)
 *
 * Compiled with -ffp-contract=off so that a*b+c is never fused, as in the reference's x86-64 build.
 *
 * How each function is pinned (tests/test_oracle_pins.py unless stated; tests/test_poses.py for the poses):
 *   - index/frequency/bounds: numpy's fftfreq and the worked examples of Frequency.hpp:34-35, :57-58, :82-83,
 *     :117-118, :253-256.
 *   - poses (noa_amd.xform.euler2matrix, the workloads' rotations): scipy's independent Rotation for the 12 axis
 *     sequences, intrinsic and extrinsic, both handednesses; a round trip through scipy's inverse;
 *     rotate_z(90deg) maps x to y; quaternion == matrix; orthonormality.
 *   - windowed sinc: numpy's sinc times scipy's Blackman window, w(0)=1, the window-size examples of
 *     FourierProject.hpp:118-122, :140-141.
 *   - interpolate_spectrum: exact copies at integer frequencies, exact reproduction of a linear Hermitian field
 *     inside the bounds, the flip and out-of-bounds rules, NEAREST without the conjugate; the windows of 4-8 taps
 *     against numpy's textbook kernels.
 *   - rasterize: identity-pose closed form incl. the x=0 Hermitian duplicate, the adjoint of extraction in FP64,
 *     mass conservation (closed form of the total weight), an FP64 restatement written separately in numpy
 *     (test_rasterize_fp64_brute_force_in_numpy), negation of flipped values, and the reference's remap and
 *     weights invariants (TestXformFourierInsertRasterize.cpp:72-152).
 *   - insert (interpolate): identity-pose closed form, a linear field at random poses, linearity, the EWS term
 *     against the paraboloid, matrix == quaternion, weights == values, and an FP64 sweep of every voxel one slice
 *     at a time written separately in numpy (test_insert_fp64_per_slice_scatter_in_numpy).
 *   - extract: identity-pose closed form, a linear field at random poses, the reference's remap invariant
 *     (TestXformFourierExtract.cpp:116-153), EWS/scaling/target/w-window closed forms, and an FP64 restatement
 *     written separately in numpy (test_extract_fp64_restatement_in_numpy).
 *   - insert's per-voxel sums at a voxel list (oracle_insert_central_slices_3d_voxels) run the same operator
 *     object as the whole-volume entry point; they agree bit for bit on whole volumes.
 *
 * One deviation from noa, shared with the HIP kernels: rasterisation skips taps at x > n/2. noa bounds x by the
 * logical width there (FourierProject.hpp:330-332) and, for a contiguous array, writes such taps into the next row;
 * their weight is zero for cutoffs <= 0.5 and target shapes equal to the grid.
 */
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <algorithm>
#include <limits>
#include <type_traits>
#include <omp.h>

#include "../include/noa_hip.h"

namespace oracle {

using i64 = int64_t;

// ---------------------------------------------------------------------------------------------------------
// Value types. Complex arithmetic follows src/noa/base/Complex.hpp:157-216.
// ---------------------------------------------------------------------------------------------------------
template<class T> struct Cx { T re, im; };
template<class T> struct is_cx { static constexpr bool value = false; };
template<class T> struct is_cx<Cx<T>> { static constexpr bool value = true; };
template<class T> struct real_of { using type = T; };
template<class T> struct real_of<Cx<T>> { using type = T; };

template<class T> inline Cx<T> operator+(Cx<T> a, Cx<T> b) { return {a.re + b.re, a.im + b.im}; }
template<class T> inline Cx<T>& operator+=(Cx<T>& a, Cx<T> b) { a = a + b; return a; }
// Complex * real = {s*re, s*im} (Complex.hpp:203-216).
template<class T> inline Cx<T> operator*(Cx<T> a, T s) { return {s * a.re, s * a.im}; }

template<class T> inline T zero() { return T{}; }

// static_cast between value types of the same kind (real->real, complex->complex).
template<class To, class From>
inline To vcast(const From& v) {
    if constexpr (is_cx<To>::value && is_cx<From>::value) {
        using R = typename real_of<To>::type;
        return To{static_cast<R>(v.re), static_cast<R>(v.im)};
    } else if constexpr (!is_cx<To>::value && !is_cx<From>::value) {
        return static_cast<To>(v);
    } else {
        static_assert(!is_cx<To>::value, "real to complex is not a spectrum type");
        // complex -> real: cast_or_abs_squared (base/Math.hpp:220-225, Complex.hpp:398-402)
        return static_cast<To>(v.re * v.re + v.im * v.im);
    }
}

// cast_or_abs_squared<T>(value) (base/Math.hpp:220-225).
template<class To, class From>
inline To cast_or_abs_squared(const From& v) { return vcast<To>(v); }

// ---------------------------------------------------------------------------------------------------------
// Small vectors and matrices (row-major; Mat*Vec is a row-wise dot, base/Mat.hpp:419-424; dot is a left fold,
// base/Vec.hpp:846-855).
// ---------------------------------------------------------------------------------------------------------
template<class T> struct V2 { T v[2]; T& operator[](int i) { return v[i]; } const T& operator[](int i) const { return v[i]; } };
template<class T> struct V3 { T v[3]; T& operator[](int i) { return v[i]; } const T& operator[](int i) const { return v[i]; } };
template<class T> struct M22 { T m[2][2]; };
template<class T> struct M33 { T m[3][3]; };
template<class T> struct Quat { T z, y, x, w; };

template<class T> inline T dot3(const V3<T>& a, const V3<T>& b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
template<class T> inline T dot2(const V2<T>& a, const V2<T>& b) { return a[0] * b[0] + a[1] * b[1]; }

template<class T> inline V3<T> matvec(const M33<T>& r, const V3<T>& x) {
    V3<T> out;
    for (int i = 0; i < 3; ++i)
        out[i] = (r.m[i][0] * x[0] + r.m[i][1] * x[1]) + r.m[i][2] * x[2];
    return out;
}
template<class T> inline V2<T> matvec(const M22<T>& s, const V2<T>& x) {
    V2<T> out;
    for (int i = 0; i < 2; ++i)
        out[i] = s.m[i][0] * x[0] + s.m[i][1] * x[1];
    return out;
}
// Quaternion::rotate (xform/core/Quaternion.hpp:167-181).
template<class T> inline V3<T> quat_rotate(const Quat<T>& q, const V3<T>& zyx) {
    const T a0 = q.w * zyx[2] + q.y * zyx[0] - q.z * zyx[1];
    const T a1 = q.w * zyx[1] + q.z * zyx[2] - q.x * zyx[0];
    const T a2 = q.w * zyx[0] + q.x * zyx[1] - q.y * zyx[2];
    const T two = 2;
    V3<T> out;
    out[0] = zyx[0] + two * (q.x * a1 - q.y * a0);
    out[1] = zyx[1] + two * (q.z * a0 - q.x * a2);
    out[2] = zyx[2] + two * (q.y * a2 - q.z * a1);
    return out;
}

// A batched rotation (Mat33 or Quaternion) or scaling (Mat22), as BatchedParameter (runtime/core/Utils.hpp:391-413).
template<class T>
struct Rotation {
    const unsigned char* data{};
    int kind{};
    bool batched{};
    V3<T> apply(i64 b, const V3<T>& x) const {
        const i64 i = batched ? b : 0;
        if (kind == NOA_HIP_MAT33_F32 || kind == NOA_HIP_MAT33_F64) {
            M33<T> r;
            const bool f32 = kind == NOA_HIP_MAT33_F32;
            const size_t stride = f32 ? 48 : 80;   // alignas(16) Mat33 (base/Mat.hpp:177)
            const unsigned char* p = data + i * stride;
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) {
                    if (f32) { float e; std::memcpy(&e, p + (a * 3 + c) * 4, 4); r.m[a][c] = static_cast<T>(e); }
                    else { double e; std::memcpy(&e, p + (a * 3 + c) * 8, 8); r.m[a][c] = static_cast<T>(e); }
                }
            return matvec(r, x);
        } else {
            Quat<T> q;
            if (kind == NOA_HIP_QUAT_F32) {
                float e[4]; std::memcpy(e, data + i * 16, 16);
                q = {static_cast<T>(e[0]), static_cast<T>(e[1]), static_cast<T>(e[2]), static_cast<T>(e[3])};
            } else {
                double e[4]; std::memcpy(e, data + i * 32, 32);
                q = {static_cast<T>(e[0]), static_cast<T>(e[1]), static_cast<T>(e[2]), static_cast<T>(e[3])};
            }
            return quat_rotate(q, x);
        }
    }
};

template<class T>
struct Scaling {
    const unsigned char* data{};
    int kind{};      // NOA_HIP_XFORM_NONE = Empty (identity without arithmetic)
    bool batched{};
    V2<T> apply(i64 b, const V2<T>& x) const {
        if (kind == NOA_HIP_XFORM_NONE)
            return x;
        const i64 i = batched ? b : 0;
        M22<T> s;
        const bool f32 = kind == NOA_HIP_MAT22_F32;
        const unsigned char* p = data + i * (f32 ? 16 : 32);
        for (int a = 0; a < 2; ++a)
            for (int c = 0; c < 2; ++c) {
                if (f32) { float e; std::memcpy(&e, p + (a * 2 + c) * 4, 4); s.m[a][c] = static_cast<T>(e); }
                else { double e; std::memcpy(&e, p + (a * 2 + c) * 8, 8); s.m[a][c] = static_cast<T>(e); }
            }
        return matvec(s, x);
    }
};

// ---------------------------------------------------------------------------------------------------------
// Index <-> frequency (fft/core/Frequency.hpp).
// ---------------------------------------------------------------------------------------------------------
inline i64 index2frequency(bool centered, i64 index, i64 size) {               // Frequency.hpp:80-87
    if (centered) return index - size / 2;
    return index < (size + 1) / 2 ? index : index - size;
}
inline i64 frequency2index(bool centered, i64 frequency, i64 size) {           // Frequency.hpp:115-122
    if (centered) return frequency + size / 2;
    return frequency < 0 ? frequency + size : frequency;
}

// ---------------------------------------------------------------------------------------------------------
// Windowed sinc (FourierProject.hpp:97-159, noa::sinc base/Math.hpp:28-29).
// ---------------------------------------------------------------------------------------------------------
template<class T> inline T sinc(T x) { return x == 0 ? T(1) : std::sin(x) / x; }

template<class T>
inline T windowed_sinc(T fftfreq, T fftfreq_sinc, T fftfreq_blackman) {
    const T PI = static_cast<T>(3.141592653589793238462643383279502884);
    fftfreq *= PI;
    const T s = sinc(fftfreq / fftfreq_sinc);
    const T blackman_cutoff = fftfreq / fftfreq_blackman;
    const T blackman = static_cast<T>(0.42) + static_cast<T>(0.5) * std::cos(blackman_cutoff) +
                       static_cast<T>(0.08) * std::cos(2 * blackman_cutoff);
    return s * blackman;
}

// noa::allclose(a, b) with the default epsilon 1e-6 (base/Math.hpp allclose: |a-b| <= eps).
inline bool allclose_default(double a, double b) { return std::abs(a - b) <= 1e-6; }

template<class T>
inline i64 blackman_window_size(T fftfreq_blackman, T spectrum_size) {  // FourierProject.hpp:114-137
    double samples = static_cast<double>(spectrum_size) * static_cast<double>(fftfreq_blackman);
    if (allclose_default(samples, 1.))
        samples = std::round(samples);
    else
        samples = std::ceil(samples);
    const i64 n = std::max(i64{1}, static_cast<i64>(samples));
    return 2 * n + 1 - 2;
}

template<class T>
inline T w_index_to_fftfreq_offset(i64 w, i64 window_size, T spectrum_size) {  // :142-145
    return static_cast<T>(w - window_size / 2) / spectrum_size;
}

// ---------------------------------------------------------------------------------------------------------
// Strided readers/writers over host memory, typed by template. A reader can also be a constant (AccessorValue)
// or empty.
// ---------------------------------------------------------------------------------------------------------
template<class T>
struct Reader3 {   // (batch, j, k) view, or (batch, i, j, k) for the 3D spectrum
    const T* p{};
    i64 s0{}, s1{}, s2{}, s3{};
    bool is_value{};
    T value{};
    T operator()(i64 a, i64 b, i64 c) const { return is_value ? value : p[a * s0 + b * s1 + c * s2]; }
    T operator()(i64 a, i64 b, i64 c, i64 d) const { return is_value ? value : p[a * s0 + b * s1 + c * s2 + d * s3]; }
};

template<class T>
struct Writer3 {
    T* p{};
    i64 s0{}, s1{}, s2{};
    T& operator()(i64 a, i64 b, i64 c) const { return p[a * s0 + b * s1 + c * s2]; }
};

// ---------------------------------------------------------------------------------------------------------
// interpolate_spectrum<REMAP, INTERP, N> for N = 2, 3 (xform/core/Interpolation.hpp:435-526).
// Supports NEAREST and LINEAR (window size 2: early flip, :459-463, :522-523).
// `freq` is the centred frequency in samples; `shape` the logical shape; rfft along the last dimension.
// ---------------------------------------------------------------------------------------------------------
// interpolation_weights (Interpolation.hpp:81-141) of the windows larger than 2, for one dimension, in the value's
// real precision R: CUBIC (Keys, A = -0.5), CUBIC_BSPLINE, LANCZOS4/6/8 (OpenCV's one-sincos form, normalised).
inline int window_size(int interp) {   // Interp::window_size (Interp.hpp:55-77)
    const int base = interp >= 100 ? interp - 100 : interp;
    switch (base) {
        case NOA_HIP_INTERP_NEAREST: case NOA_HIP_INTERP_LINEAR: return 2;
        case NOA_HIP_INTERP_CUBIC: case NOA_HIP_INTERP_CUBIC_BSPLINE: case NOA_HIP_INTERP_LANCZOS4: return 4;
        case NOA_HIP_INTERP_LANCZOS6: return 6;
        default: return 8;
    }
}

template<class R>
void window_weights(int interp, R f, R* w) {
    const int base = interp >= 100 ? interp - 100 : interp;
    if (base == NOA_HIP_INTERP_CUBIC) {
        const R A = static_cast<R>(-0.5);
        w[0] = ((A * (f + 1) - 5 * A) * (f + 1) + 8 * A) * (f + 1) - 4 * A;
        w[1] = ((A + 2) * f - (A + 3)) * f * f + 1;
        w[2] = ((A + 2) * (1 - f) - (A + 3)) * (1 - f) * (1 - f) + 1;
        w[3] = 1 - w[0] - w[1] - w[2];
    } else if (base == NOA_HIP_INTERP_CUBIC_BSPLINE) {
        const R one_sixth = static_cast<R>(1) / static_cast<R>(6);
        const R two_third = static_cast<R>(2) / static_cast<R>(3);
        const R one_minus = 1 - f;
        const R one_squared = one_minus * one_minus;
        const R squared = f * f;
        w[0] = one_sixth * one_squared * one_minus;
        w[1] = two_third - static_cast<R>(0.5) * squared * (2 - f);
        w[2] = two_third - static_cast<R>(0.5) * one_squared * (2 - one_minus);
        w[3] = one_sixth * squared * f;
    } else {   // LANCZOS4, 6, 8
        const int size = window_size(interp);
        const int center = size / 2 - 1;
        const R PI = static_cast<R>(3.1415926535897932384626433832795);
        const R s45 = static_cast<R>(0.7071067811865475);
        const R cs[8][2] = {{1, 0}, {-s45, -s45}, {0, 1}, {s45, -s45}, {-1, 0}, {s45, s45}, {0, -1}, {-s45, s45}};
        const R eps = std::numeric_limits<R>::epsilon();
        for (int i = 0; i < size; ++i) w[i] = 0;
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
            const R sn = std::sin(y0), cn = std::cos(y0);
            for (int i = 0; i < size; ++i) {
                const R fij = -(f + static_cast<R>(center) - static_cast<R>(i));
                const R yj = fij * PI * static_cast<R>(0.25);
                w[i] = (cs[i][0] * sn + cs[i][1] * cn) / (yj * yj);
                sum += w[i];
            }
        }
        const R inv = 1 / sum;
        for (int i = 0; i < size; ++i) w[i] *= inv;
    }
}

// interpolate_spectrum for windows larger than 2 (Interpolation.hpp:435-526): the flip is decided per tap (a tap
// at x < 0 reads the conjugate of the mirrored element), taps outside the bounds contribute 0.
template<class V, class Coord>
V interpolate_spectrum_window_2d(const Reader3<V>& in, i64 batch, Coord fy, Coord fx, i64 ny, i64 nx, bool centered,
                                 int interp) {
    using R = typename real_of<V>::type;
    const int size = window_size(interp);
    const i64 start = -(size - 1) / 2;
    const i64 lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    const Coord fly = std::floor(fy), flx = std::floor(fx);
    const i64 iy = static_cast<i64>(fly), ix = static_cast<i64>(flx);
    R wy[8], wx[8];
    window_weights<R>(interp, static_cast<R>(fy - fly), wy);
    window_weights<R>(interp, static_cast<R>(fx - flx), wx);
    V value = zero<V>();
    for (int y = 0; y < size; ++y) {
        V vy = zero<V>();
        for (int x = 0; x < size; ++x) {
            i64 ty = iy + start + y, tx = ix + start + x;
            R conj = 1;
            if (tx < 0) { ty = -ty; tx = -tx; conj = -1; }
            if (ty < lo_y || ty > hi_y || tx > hi_x) continue;
            V v = vcast<V>(in(batch, frequency2index(centered, ty, ny), tx));
            if constexpr (is_cx<V>::value) v.im *= conj;
            vy += v * wx[x];
        }
        value += vy * wy[y];
    }
    return value;
}

template<class V, class Coord>
V interpolate_spectrum_window_3d(const Reader3<V>& in, Coord fz, Coord fy, Coord fx, i64 nz, i64 ny, i64 nx,
                                 bool centered, int interp) {
    using R = typename real_of<V>::type;
    const int size = window_size(interp);
    const i64 start = -(size - 1) / 2;
    const i64 lo_z = -nz / 2, hi_z = (nz - 1) / 2, lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    const Coord flz = std::floor(fz), fly = std::floor(fy), flx = std::floor(fx);
    const i64 iz = static_cast<i64>(flz), iy = static_cast<i64>(fly), ix = static_cast<i64>(flx);
    R wz[8], wy[8], wx[8];
    window_weights<R>(interp, static_cast<R>(fz - flz), wz);
    window_weights<R>(interp, static_cast<R>(fy - fly), wy);
    window_weights<R>(interp, static_cast<R>(fx - flx), wx);
    V value = zero<V>();
    for (int z = 0; z < size; ++z) {
        V vz = zero<V>();
        for (int y = 0; y < size; ++y) {
            V vy = zero<V>();
            for (int x = 0; x < size; ++x) {
                i64 tz = iz + start + z, ty = iy + start + y, tx = ix + start + x;
                R conj = 1;
                if (tx < 0) { tz = -tz; ty = -ty; tx = -tx; conj = -1; }
                if (tz < lo_z || tz > hi_z || ty < lo_y || ty > hi_y || tx > hi_x) continue;
                V v = vcast<V>(in(0, frequency2index(centered, tz, nz), frequency2index(centered, ty, ny), tx));
                if constexpr (is_cx<V>::value) v.im *= conj;
                vy += v * wx[x];
            }
            vz += vy * wy[y];
        }
        value += vz * wz[z];
    }
    return value;
}

template<class V, class Coord>
V interpolate_spectrum_2d(const Reader3<V>& in, i64 batch, Coord fy, Coord fx, i64 ny, i64 nx,
                          bool centered, int interp) {
    if (window_size(interp) > 2)
        return interpolate_spectrum_window_2d<V, Coord>(in, batch, fy, fx, ny, nx, centered, interp);
    using R = typename real_of<V>::type;
    R conjugate = 1;
    if (fx < 0) { fy *= -1; fx *= -1; conjugate = -1; }    // flip_frequency (:417-426), Vec *= -1
    const i64 lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;   // frequency_bounds<RFFT> (Frequency.hpp:258-276)
    auto fetch = [&](i64 y, i64 x, R w, V& out) {
        if (y < lo_y || y > hi_y || x < 0 || x > hi_x) return;
        out += vcast<V>(in(batch, frequency2index(centered, y, ny), x)) * w;
    };
    V value = zero<V>();
    const int base = interp >= 100 ? interp - 100 : interp;
    if (base == NOA_HIP_INTERP_NEAREST) {
        // The NEAREST branch returns before the conjugation of the early flip (:479-484): after a flip the
        // imaginary part is NOT negated. Reproduced as is.
        fetch(static_cast<i64>(std::round(fy)), static_cast<i64>(std::round(fx)), R(1), value);
        return value;
    }
    const Coord fly = std::floor(fy), flx = std::floor(fx);
    const i64 iy = static_cast<i64>(fly), ix = static_cast<i64>(flx);
    const R ty = static_cast<R>(fy - fly), tx = static_cast<R>(fx - flx);
    const R wy[2] = {1 - ty, ty}, wx[2] = {1 - tx, tx};
    for (int y = 0; y < 2; ++y) {
        V vy = zero<V>();
        for (int x = 0; x < 2; ++x)
            fetch(iy + y, ix + x, wx[x], vy);
        value += vy * wy[y];
    }
    if constexpr (is_cx<V>::value) value.im *= conjugate;
    return value;
}

template<class V, class Coord>
V interpolate_spectrum_3d(const Reader3<V>& in, Coord fz, Coord fy, Coord fx, i64 nz, i64 ny, i64 nx,
                          bool centered, int interp) {
    if (window_size(interp) > 2)
        return interpolate_spectrum_window_3d<V, Coord>(in, fz, fy, fx, nz, ny, nx, centered, interp);
    using R = typename real_of<V>::type;
    R conjugate = 1;
    if (fx < 0) { fz *= -1; fy *= -1; fx *= -1; conjugate = -1; }
    const i64 lo_z = -nz / 2, hi_z = (nz - 1) / 2, lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    auto fetch = [&](i64 z, i64 y, i64 x, R w, V& out) {
        if (z < lo_z || z > hi_z || y < lo_y || y > hi_y || x < 0 || x > hi_x) return;
        out += vcast<V>(in(0, frequency2index(centered, z, nz), frequency2index(centered, y, ny), x)) * w;
    };
    V value = zero<V>();
    const int base = interp >= 100 ? interp - 100 : interp;
    if (base == NOA_HIP_INTERP_NEAREST) {
        fetch(static_cast<i64>(std::round(fz)), static_cast<i64>(std::round(fy)),
              static_cast<i64>(std::round(fx)), R(1), value);
        return value;   // no conjugation after the flip, as in the reference (:479-484)
    }
    const Coord flz = std::floor(fz), fly = std::floor(fy), flx = std::floor(fx);
    const i64 iz = static_cast<i64>(flz), iy = static_cast<i64>(fly), ix = static_cast<i64>(flx);
    const R tz = static_cast<R>(fz - flz), ty = static_cast<R>(fy - fly), tx = static_cast<R>(fx - flx);
    const R wz[2] = {1 - tz, tz}, wy[2] = {1 - ty, ty}, wx[2] = {1 - tx, tx};
    for (int z = 0; z < 2; ++z) {
        V vz = zero<V>();
        for (int y = 0; y < 2; ++y) {
            V vy = zero<V>();
            for (int x = 0; x < 2; ++x)
                fetch(iz + z, iy + y, ix + x, wx[x], vy);
            vz += vy * wy[y];
        }
        value += vz * wz[z];
    }
    if constexpr (is_cx<V>::value) value.im *= conjugate;
    return value;
}

// ---------------------------------------------------------------------------------------------------------
// Execution: the CPU iwise (runtime/cpu/Iwise.hpp:97-117). Parallel only above 2^20 indices, with
// threads = min(n_threads, n / 2^20). The rasterisation's complex atomic add is two `omp atomic` adds
// (cpu/ComputeHandle.hpp:39-55); serial mode uses a plain +=.
// ---------------------------------------------------------------------------------------------------------
inline int iwise_threads(i64 n_elements, int n_threads) {
    const i64 per = i64{1} << 20;
    int t = n_elements <= per ? 1 : n_threads;
    if (t > 1) t = static_cast<int>(std::min<i64>(n_threads, n_elements / per));
    return std::max(t, 1);
}

template<class Op>
void iwise3(i64 n0, i64 n1, i64 n2, const Op& op, int n_threads, bool force_threads) {
    const int t = force_threads ? std::max(1, n_threads) : iwise_threads(n0 * n1 * n2, n_threads);
    if (t > 1) {
        #pragma omp parallel num_threads(t)
        {
            Op local = op;   // firstprivate(op)
            #pragma omp for collapse(3)
            for (i64 i = 0; i < n0; ++i)
                for (i64 j = 0; j < n1; ++j)
                    for (i64 k = 0; k < n2; ++k)
                        local(i, j, k, true);
        }
    } else {
        for (i64 i = 0; i < n0; ++i)
            for (i64 j = 0; j < n1; ++j)
                for (i64 k = 0; k < n2; ++k)
                    op(i, j, k, false);
    }
}

template<class Op>
void iwise4(i64 n0, i64 n1, i64 n2, i64 n3, const Op& op, int n_threads, bool force_threads) {
    const int t = force_threads ? std::max(1, n_threads) : iwise_threads(n0 * n1 * n2 * n3, n_threads);
    if (t > 1) {
        #pragma omp parallel num_threads(t)
        {
            Op local = op;
            #pragma omp for collapse(4)
            for (i64 i = 0; i < n0; ++i)
                for (i64 j = 0; j < n1; ++j)
                    for (i64 k = 0; k < n2; ++k)
                        for (i64 l = 0; l < n3; ++l)
                            local(i, j, k, l, true);
        }
    } else {
        for (i64 i = 0; i < n0; ++i)
            for (i64 j = 0; j < n1; ++j)
                for (i64 k = 0; k < n2; ++k)
                    for (i64 l = 0; l < n3; ++l)
                        op(i, j, k, l, false);
    }
}

template<class T>
inline void atomic_add(T& dst, const T& v, bool parallel) {
    if constexpr (is_cx<T>::value) {
        if (parallel) {
            #pragma omp atomic
            dst.re += v.re;
            #pragma omp atomic
            dst.im += v.im;
        } else {
            dst = dst + v;
        }
    } else {
        if (parallel) {
            #pragma omp atomic
            dst += v;
        } else {
            dst += v;
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Common per-call parameters.
// ---------------------------------------------------------------------------------------------------------
template<class Coord>
struct Geometry {
    Rotation<Coord> rot;
    Scaling<Coord> scale;
    bool general{};           // scale or EWS present (FourierProject.hpp:1312-1316)
    Coord ews_inv[2]{};       // 1 / (2 * ews_radius) per component, or {0, 0}

    // fourier_slice2grid (FourierProject.hpp:41-68).
    V3<Coord> slice2grid(V2<Coord> f, i64 b) const {
        V3<Coord> f3;
        if (general) {
            f = scale.apply(b, f);
            f3[0] = (ews_inv[0] * f[0]) * f[0] + (ews_inv[1] * f[1]) * f[1];
        } else {
            f3[0] = 0;
        }
        f3[1] = f[0];
        f3[2] = f[1];
        return rot.apply(b, f3);
    }
    // fourier_grid2slice (FourierProject.hpp:71-95).
    void grid2slice(const V3<Coord>& f, i64 b, Coord& fz, V2<Coord>& f2) const {
        const V3<Coord> r = rot.apply(b, f);
        f2[0] = r[1];
        f2[1] = r[2];
        fz = r[0];
        if (general) {
            fz -= (ews_inv[0] * f2[0]) * f2[0] + (ews_inv[1] * f2[1]) * f2[1];
            f2 = scale.apply(b, f2);
        }
    }
};

// fourier_project_has_scale: a single Mat22 counts only if it differs from the identity (:1234-1244).
inline bool has_scale(const noa_hip_transform& scaling) {
    bool has = scaling.kind != NOA_HIP_XFORM_NONE;
    if (has && !scaling.batched) {
        double e[4];
        if (scaling.kind == NOA_HIP_MAT22_F32) { float f[4]; std::memcpy(f, scaling.data, 16); for (int i = 0; i < 4; ++i) e[i] = f[i]; }
        else std::memcpy(e, scaling.data, 32);
        has = !(e[0] == 1 && e[1] == 0 && e[2] == 0 && e[3] == 1);
    }
    return has;
}

// `general` (WrapNoEwaldAndScale<false>) is decided by the caller: EWS or any scaling of the call.
template<class Coord>
Geometry<Coord> make_geometry(const noa_hip_transform& rotation, const noa_hip_transform& scaling,
                              const double ews_radius[2], bool general) {
    Geometry<Coord> g;
    g.rot.data = static_cast<const unsigned char*>(rotation.data);
    g.rot.kind = rotation.kind;
    g.rot.batched = rotation.batched != 0;
    g.scale.data = static_cast<const unsigned char*>(scaling.data);
    g.scale.kind = has_scale(scaling) ? scaling.kind : NOA_HIP_XFORM_NONE;
    g.scale.batched = scaling.batched != 0;
    g.general = general;
    if (g.general) {
        const Coord r0 = static_cast<Coord>(ews_radius[0]), r1 = static_cast<Coord>(ews_radius[1]);
        if (r0 != 0 || r1 != 0) {   // `ews_radius != 0 ? 1 / (2 * ews_radius) : {}` (:231)
            g.ews_inv[0] = 1 / (2 * r0);
            g.ews_inv[1] = 1 / (2 * r1);
        }
    }
    return g;
}

inline bool has_ews(const double ews_radius[2]) { return ews_radius[0] != 0 || ews_radius[1] != 0; }

template<class Coord>
Geometry<Coord> make_geometry(const noa_hip_project_args& a) {
    return make_geometry<Coord>(a.rotation, a.scaling, a.ews_radius, has_ews(a.ews_radius) || has_scale(a.scaling));
}

inline bool remap_src_centered(int remap) { return (remap & 4) != 0; }   // is_xc2xx
inline bool remap_dst_centered(int remap) { return (remap & 8) != 0; }   // is_xx2xc

template<class T>
Reader3<T> reader_slices(const noa_hip_view& v) {   // filter {0, 2, 3}
    Reader3<T> r;
    r.p = static_cast<const T*>(v.data);
    r.s0 = v.strides[0]; r.s1 = v.strides[2]; r.s2 = v.strides[3];
    r.is_value = v.is_value != 0;
    if constexpr (is_cx<T>::value) r.value = T{static_cast<typename real_of<T>::type>(v.value[0]), static_cast<typename real_of<T>::type>(v.value[1])};
    else r.value = static_cast<T>(v.value[0]);
    return r;
}

template<class T>
Reader3<T> reader_interp2(const noa_hip_view& v) {   // to_interpolator_input<2>: batch stride 0 if B == 1
    Reader3<T> r = reader_slices<T>(v);
    if (v.shape[0] == 1) r.s0 = 0;
    return r;
}

template<class T>
Reader3<T> reader_volume(const noa_hip_view& v) {    // 3D interpolator input (batch, z, y, x)
    Reader3<T> r;
    r.p = static_cast<const T*>(v.data);
    r.s0 = 0; r.s1 = v.strides[1]; r.s2 = v.strides[2]; r.s3 = v.strides[3];
    r.is_value = v.is_value != 0;
    return r;
}

template<class T>
Writer3<T> writer_volume(const noa_hip_view& v) {    // filter {1, 2, 3}
    Writer3<T> w;
    w.p = static_cast<T*>(v.data);
    w.s0 = v.strides[1]; w.s1 = v.strides[2]; w.s2 = v.strides[3];
    return w;
}

template<class T>
Writer3<T> writer_slices(const noa_hip_view& v) {    // filter {0, 2, 3}
    Writer3<T> w;
    w.p = static_cast<T*>(v.data);
    w.s0 = v.strides[0]; w.s1 = v.strides[2]; w.s2 = v.strides[3];
    return w;
}

inline void target3(const noa_hip_project_args& a, const int64_t* grid_logical, i64 out[3]) {
    const bool any0 = a.target_shape[0] == 0 || a.target_shape[1] == 0 || a.target_shape[2] == 0 || a.target_shape[3] == 0;
    for (int i = 0; i < 3; ++i)
        out[i] = any0 ? grid_logical[i + 1] : a.target_shape[i + 1];
}

// ---------------------------------------------------------------------------------------------------------
// FourierInsertRasterize (FourierProject.hpp:161-390).
// ---------------------------------------------------------------------------------------------------------
template<class Coord, class In, class Out, class W>
struct RasterizeOp {
    using RIn = typename real_of<In>::type;
    using ROut = typename real_of<Out>::type;
    Reader3<In> slices;
    Reader3<W> wslices; bool has_wslices{};
    Writer3<Out> volume;
    Writer3<W> wvolume; bool has_wvolume{};
    Geometry<Coord> geo;
    i64 grid[3]{};           // logical (D, H, W)
    i64 slice_h{};
    Coord f_slice[2]{};      // (H, W)
    Coord f_target[3]{};
    Coord cutoff_sqd{};
    bool slices_centered{}, volume_centered{};

    void operator()(i64 b, i64 y, i64 u, bool parallel) const {
        const i64 v = index2frequency(slices_centered, y, slice_h);
        V2<Coord> f2;
        f2[0] = static_cast<Coord>(v) / f_slice[0];
        f2[1] = static_cast<Coord>(u) / f_slice[1];
        V3<Coord> f3 = geo.slice2grid(f2, b);
        if (dot3(f3, f3) > cutoff_sqd)
            return;
        RIn conjugate = 1;
        if (f3[2] < 0) {
            f3[0] = -f3[0]; f3[1] = -f3[1]; f3[2] = -f3[2];
            if constexpr (is_cx<In>::value) conjugate = -1;
        }
        V3<Coord> freq;
        for (int i = 0; i < 3; ++i) freq[i] = f3[i] * f_target[i];

        // get_input_value_ (:270-277): complex -> complex multiplies by `conjugate` (a negation when -1).
        Out value;
        const In raw = slices(b, y, u);
        if constexpr (is_cx<In>::value && is_cx<Out>::value) value = vcast<Out>(raw * conjugate);
        else value = cast_or_abs_squared<Out>(raw);
        // get_input_weight_ (:279-286).
        W weight{};
        if (has_wvolume) weight = has_wslices ? static_cast<W>(wslices(b, y, u)) : W(1);

        rasterize(value, weight, freq, parallel);
    }

    void rasterize(Out value, W weight, const V3<Coord>& freq, bool parallel) const {
        i64 base0[3];
        for (int i = 0; i < 3; ++i) base0[i] = static_cast<i64>(std::floor(freq[i]));
        // set_rasterization_weights_ (:290-304)
        Coord fr[2][3];
        for (int i = 0; i < 3; ++i) {
            fr[1][i] = freq[i] - static_cast<Coord>(base0[i]);
            fr[0][i] = 1 - fr[1][i];
        }
        Coord k[2][2][2];
        for (int w = 0; w < 2; ++w)
            for (int v = 0; v < 2; ++v)
                for (int u = 0; u < 2; ++u)
                    k[w][v][u] = fr[w][0] * fr[v][1] * fr[u][2];

        const i64 width = grid[2] / 2 + 1;   // stored rfft width (see quirk note below)
        for (int w = 0; w < 2; ++w) {
            for (int v = 0; v < 2; ++v) {
                for (int u = 0; u < 2; ++u) {
                    const i64 iw = frequency2index(volume_centered, base0[0] + w, grid[0]);
                    const i64 iv = frequency2index(volume_centered, base0[1] + v, grid[1]);
                    const i64 iu = base0[2] + u;
                    // The reference bounds iu by the logical W (:332); a tap with iu >= W/2+1 only occurs at
                    // freq_x == W/2 with a zero weight (cutoff <= 0.5), so skipping it is identical.
                    if (iw >= 0 && iw < grid[0] && iv >= 0 && iv < grid[1] && iu >= 0 && iu < width) {
                        const Coord fraction = k[w][v][u];
                        atomic_add(volume(iw, iv, iu), value * static_cast<ROut>(fraction), parallel);
                        if (has_wvolume)
                            atomic_add(wvolume(iw, iv, iu), weight * static_cast<W>(fraction), parallel);
                    }
                }
            }
        }
        // Hermitian x=0 plane (:349-372).
        if (base0[2] == 0) {
            if constexpr (is_cx<Out>::value) value.im = -value.im;
            for (int w = 0; w < 2; ++w) {
                for (int v = 0; v < 2; ++v) {
                    const i64 iw = frequency2index(volume_centered, -(base0[0] + w), grid[0]);
                    const i64 iv = frequency2index(volume_centered, -(base0[1] + v), grid[1]);
                    if (iw >= 0 && iw < grid[0] && iv >= 0 && iv < grid[1]) {
                        const Coord fraction = k[w][v][0];
                        atomic_add(volume(iw, iv, 0), value * static_cast<ROut>(fraction), parallel);
                        if (has_wvolume)
                            atomic_add(wvolume(iw, iv, 0), weight * static_cast<W>(fraction), parallel);
                    }
                }
            }
        }
    }
};

// ---------------------------------------------------------------------------------------------------------
// FourierInsertInterpolate (FourierProject.hpp:392-539).
// ---------------------------------------------------------------------------------------------------------
template<class Coord, class In, class Out, class W>
struct InsertOp {
    Reader3<In> slices;
    Reader3<W> wslices; bool has_wslices{};
    Writer3<Out> volume;
    Writer3<W> wvolume; bool has_wvolume{};
    Geometry<Coord> geo;
    i64 grid[3]{};
    i64 n_slices{};
    i64 slice_shape[2]{};
    Coord f_slice[2]{};
    Coord f_target[3]{};
    Coord cutoff_sqd{}, fftfreq_sinc{}, fftfreq_blackman{};
    bool slices_centered{}, volume_centered{};
    int interp{};

    void operator()(i64 oz, i64 oy, i64 ox, bool) const {
        In value = zero<In>();
        W weights{};
        if (!sum_at(oz, oy, ox, value, weights))
            return;
        Out& out = volume(oz, oy, ox);
        out = add(out, cast_or_abs_squared<Out>(value));
        if (has_wvolume) wvolume(oz, oy, ox) += weights;
    }

    // The sums over the slices of one voxel (:478-515); false (and nothing summed) outside the cutoff.
    bool sum_at(i64 oz, i64 oy, i64 ox, In& value, W& weights) const {
        V3<Coord> f;
        f[0] = static_cast<Coord>(index2frequency(volume_centered, oz, grid[0])) / f_target[0];
        f[1] = static_cast<Coord>(index2frequency(volume_centered, oy, grid[1])) / f_target[1];
        f[2] = static_cast<Coord>(ox) / f_target[2];
        if (dot3(f, f) > cutoff_sqd)
            return false;
        value = zero<In>();
        weights = W{};
        for (i64 i = 0; i < n_slices; ++i) {
            Coord fz;
            V2<Coord> f2;
            geo.grid2slice(f, i, fz, f2);
            In i_value = zero<In>();
            W i_weights{};
            if (std::abs(fz) <= fftfreq_blackman) {
                const Coord window = windowed_sinc(fz, fftfreq_sinc, fftfreq_blackman);
                const Coord py = f2[0] * f_slice[0], px = f2[1] * f_slice[1];
                using RIn = typename real_of<In>::type;
                i_value = mul(interpolate_spectrum_2d<In, Coord>(slices, i, py, px, slice_shape[0], slice_shape[1],
                                                                 slices_centered, interp),
                              static_cast<RIn>(window));
                if (has_wvolume) {
                    if (has_wslices)
                        i_weights = interpolate_spectrum_2d<W, Coord>(wslices, i, py, px, slice_shape[0], slice_shape[1],
                                                                      slices_centered, interp) * static_cast<W>(window);
                    else
                        i_weights = static_cast<W>(window);
                }
            }
            value = add(value, i_value);
            if (has_wvolume) weights += i_weights;
        }
        return true;
    }

    template<class T, class R> static T mul(T a, R s) { if constexpr (is_cx<T>::value) return a * s; else return a * s; }
    template<class T> static T add(T a, T b) { if constexpr (is_cx<T>::value) return a + b; else return a + b; }
};

// ---------------------------------------------------------------------------------------------------------
// FourierExtract (FourierProject.hpp:541-725).
// ---------------------------------------------------------------------------------------------------------
template<class Coord, class In, class Out, class W>
struct ExtractOp {
    using ROut = typename real_of<Out>::type;
    Reader3<In> volume;
    Reader3<W> wvolume; bool has_wvolume{};
    Writer3<Out> slices;
    Writer3<W> wslices; bool has_wslices{};
    Geometry<Coord> geo;
    i64 grid[3]{};
    i64 slice_h{};
    Coord f_slice[2]{};
    Coord f_target[3]{};
    Coord cutoff_sqd{}, fftfreq_sinc{}, fftfreq_blackman{};
    i64 blackman_size{};
    bool volume_centered{}, slices_centered{};
    int interp{};

    V3<Coord> fftfreq_in_volume(i64 b, i64 y, i64 x) const {   // compute_fftfreq_in_volume_ (:699-705)
        V2<Coord> f2;
        f2[0] = static_cast<Coord>(index2frequency(slices_centered, y, slice_h)) / f_slice[0];
        f2[1] = static_cast<Coord>(x) / f_slice[1];
        return geo.slice2grid(f2, b);
    }

    In sample(const V3<Coord>& f) const {
        return interpolate_spectrum_3d<In, Coord>(volume, f[0] * f_target[0], f[1] * f_target[1], f[2] * f_target[2],
                                                  grid[0], grid[1], grid[2], volume_centered, interp);
    }
    W sample_weight(const V3<Coord>& f) const {
        return interpolate_spectrum_3d<W, Coord>(wvolume, f[0] * f_target[0], f[1] * f_target[1], f[2] * f_target[2],
                                                 grid[0], grid[1], grid[2], volume_centered, interp);
    }

    // 3D operator (:632-657).
    void operator()(i64 b, i64 y, i64 x, bool) const {
        const V3<Coord> f = fftfreq_in_volume(b, y, x);
        Out value = zero<Out>();
        W weight{};
        if (dot3(f, f) <= cutoff_sqd) {
            value = cast_or_abs_squared<Out>(sample(f));
            if (has_wslices)
                weight = has_wvolume ? static_cast<W>(sample_weight(f)) : W(1);
        }
        slices(b, y, x) = value;
        if (has_wslices) wslices(b, y, x) = weight;
    }

    // 4D operator (:661-694).
    void operator()(i64 b, i64 w, i64 y, i64 x, bool parallel) const {
        V3<Coord> f = fftfreq_in_volume(b, y, x);
        const Coord offset = w_index_to_fftfreq_offset<Coord>(w, blackman_size, f_target[0]);
        f[0] += offset;
        if (dot3(f, f) > cutoff_sqd)
            return;
        const Coord conv = windowed_sinc(offset, fftfreq_sinc, fftfreq_blackman);
        const Out value = cast_or_abs_squared<Out>(sample(f));
        atomic_add(slices(b, y, x), value * static_cast<ROut>(conv), parallel);
        if (has_wslices) {
            const W weight = has_wvolume ? static_cast<W>(sample_weight(f)) : W(1);
            atomic_add(wslices(b, y, x), weight * static_cast<W>(conv), parallel);
        }
    }
};

// ---------------------------------------------------------------------------------------------------------
// FourierInsertExtract (FourierProject.hpp:727-1006): extraction from a virtual volume made of central slices.
// ---------------------------------------------------------------------------------------------------------
template<class Coord, class In, class Out, class W>
struct InsertExtractOp {
    using RIn = typename real_of<In>::type;
    using ROut = typename real_of<Out>::type;
    Reader3<In> slices;
    Reader3<W> wslices; bool has_wslices{};
    Writer3<Out> out;
    Writer3<W> wout; bool has_wout{};
    Geometry<Coord> in_geo, out_geo;     // insert: fwd scaling + inv rotation; extract: inv scaling + fwd rotation
    i64 n_in{}, in_shape[2]{}, out_h{};
    Coord f_in[2]{}, f_out[2]{};
    Coord volume_z{}, cutoff_sqd{}, in_sinc{}, in_blackman{}, w_sinc{}, w_blackman{};
    i64 w_size{};
    bool in_centered{}, out_centered{}, add{}, correct{};
    int interp{};

    V3<Coord> fftfreq_in_volume(i64 b, i64 y, i64 x) const {   // compute_fftfreq_in_volume_ (:927-933)
        V2<Coord> f2;
        f2[0] = static_cast<Coord>(index2frequency(out_centered, y, out_h)) / f_out[0];
        f2[1] = static_cast<Coord>(x) / f_out[1];
        return out_geo.slice2grid(f2, b);
    }

    // sample_virtual_volume_ (:935-977): every input slice within the window along its normal contributes its
    // interpolated value (and weight, or 1) times the windowed sinc, in slice order.
    void sample(const V3<Coord>& f, bool correct_weights, In& value, W& weight) const {
        value = zero<In>();
        weight = W{};
        for (i64 i = 0; i < n_in; ++i) {
            Coord fz;
            V2<Coord> f2;
            in_geo.grid2slice(f, i, fz, f2);
            if (std::abs(fz) < in_blackman) {
                const Coord window = windowed_sinc(fz, in_sinc, in_blackman);
                const Coord py = f2[0] * f_in[0], px = f2[1] * f_in[1];
                value += interpolate_spectrum_2d<In, Coord>(slices, i, py, px, in_shape[0], in_shape[1], in_centered,
                                                            interp) * static_cast<RIn>(window);
                if (has_wslices)
                    weight += interpolate_spectrum_2d<W, Coord>(wslices, i, py, px, in_shape[0], in_shape[1],
                                                                in_centered, interp) * static_cast<W>(window);
                else
                    weight += static_cast<W>(window);
            }
        }
        if (correct_weights) {
            const W final_weight = std::max(W(1), std::abs(weight));
            if constexpr (is_cx<In>::value) value = In{value.re / static_cast<RIn>(final_weight), value.im / static_cast<RIn>(final_weight)};
            else value = value / static_cast<RIn>(final_weight);
        }
    }

    // 3D operator (:858-886).
    void operator()(i64 b, i64 y, i64 x, bool) const {
        const V3<Coord> f = fftfreq_in_volume(b, y, x);
        if (dot3(f, f) > cutoff_sqd) {
            if (!add) {
                out(b, y, x) = zero<Out>();
                if (has_wout) wout(b, y, x) = W{};
            }
            return;
        }
        In value;
        W weight;
        sample(f, correct, value, weight);
        Out& o = out(b, y, x);
        o = add ? o + cast_or_abs_squared<Out>(value) : cast_or_abs_squared<Out>(value);
        if (has_wout) {
            W& w = wout(b, y, x);
            w = add ? w + weight : weight;
        }
    }

    // 4D operator (:890-924): one w offset of the windowed sum along the depth of the virtual volume.
    void operator()(i64 b, i64 w, i64 y, i64 x, bool parallel) const {
        V3<Coord> f = fftfreq_in_volume(b, y, x);
        const Coord offset = w_index_to_fftfreq_offset<Coord>(w, w_size, volume_z);
        f[0] += offset;
        if (dot3(f, f) > cutoff_sqd)
            return;
        In value;
        W weight;
        sample(f, false, value, weight);
        const Coord conv = windowed_sinc(offset, w_sinc, w_blackman);
        atomic_add(out(b, y, x), cast_or_abs_squared<Out>(value) * static_cast<ROut>(conv), parallel);
        if (has_wout)
            atomic_add(wout(b, y, x), weight * static_cast<W>(conv), parallel);
    }
};

// ---------------------------------------------------------------------------------------------------------
// Entry points.
// ---------------------------------------------------------------------------------------------------------
thread_local std::string g_error;

template<class Coord, class In, class Out, class W>
int run_rasterize(const noa_hip_project_args& a, int n_threads, bool force) {
    RasterizeOp<Coord, In, Out, W> op;
    op.slices = reader_slices<In>(a.input);
    op.has_wslices = a.input_weight.data != nullptr || a.input_weight.is_value;
    if (op.has_wslices) op.wslices = reader_slices<W>(a.input_weight);
    op.volume = writer_volume<Out>(a.output);
    op.has_wvolume = a.output_weight.data != nullptr;
    if (op.has_wvolume) op.wvolume = writer_volume<W>(a.output_weight);
    op.geo = make_geometry<Coord>(a);
    for (int i = 0; i < 3; ++i) op.grid[i] = a.output_shape[i + 1];
    op.slice_h = a.input_shape[2];
    op.f_slice[0] = static_cast<Coord>(a.input_shape[2]);
    op.f_slice[1] = static_cast<Coord>(a.input_shape[3]);
    i64 t[3]; target3(a, a.output_shape, t);
    for (int i = 0; i < 3; ++i) op.f_target[i] = static_cast<Coord>(t[i]);
    op.cutoff_sqd = std::max(static_cast<Coord>(a.fftfreq_cutoff), Coord(0));
    op.cutoff_sqd *= op.cutoff_sqd;
    op.slices_centered = remap_src_centered(a.remap);
    op.volume_centered = remap_dst_centered(a.remap);
    iwise3(a.input_shape[0], a.input_shape[2], a.input_shape[3] / 2 + 1, op, n_threads, force);
    return 0;
}

template<class Coord, class In, class Out, class W>
InsertOp<Coord, In, Out, W> make_insert_op(const noa_hip_project_args& a) {
    InsertOp<Coord, In, Out, W> op;
    op.slices = reader_interp2<In>(a.input);
    op.has_wslices = a.input_weight.data != nullptr || a.input_weight.is_value;
    if (op.has_wslices) op.wslices = reader_interp2<W>(a.input_weight);
    op.volume = writer_volume<Out>(a.output);
    op.has_wvolume = a.output_weight.data != nullptr;
    if (op.has_wvolume) op.wvolume = writer_volume<W>(a.output_weight);
    op.geo = make_geometry<Coord>(a);
    for (int i = 0; i < 3; ++i) op.grid[i] = a.output_shape[i + 1];
    op.n_slices = a.input_shape[0];
    op.slice_shape[0] = a.input_shape[2];
    op.slice_shape[1] = a.input_shape[3];
    op.f_slice[0] = static_cast<Coord>(a.input_shape[2]);
    op.f_slice[1] = static_cast<Coord>(a.input_shape[3]);
    i64 t[3]; target3(a, a.output_shape, t);
    for (int i = 0; i < 3; ++i) op.f_target[i] = static_cast<Coord>(t[i]);
    op.cutoff_sqd = std::max(static_cast<Coord>(a.fftfreq_cutoff), Coord(0));
    op.cutoff_sqd *= op.cutoff_sqd;
    const Coord max_output_size = static_cast<Coord>(std::min({t[0], t[1], t[2]}));
    op.fftfreq_sinc = std::max(static_cast<Coord>(a.fftfreq_sinc), 1 / max_output_size);
    op.fftfreq_blackman = std::max(static_cast<Coord>(a.fftfreq_blackman), 1 / max_output_size);
    op.slices_centered = remap_src_centered(a.remap);
    op.volume_centered = remap_dst_centered(a.remap);
    op.interp = a.interp;
    return op;
}

template<class Coord, class In, class Out, class W>
int run_insert(const noa_hip_project_args& a, int n_threads, bool force) {
    const auto op = make_insert_op<Coord, In, Out, W>(a);
    iwise3(a.output_shape[1], a.output_shape[2], a.output_shape[3] / 2 + 1, op, n_threads, force);
    return 0;
}

// The insertion's per-voxel sums (value and weight, before they are added to the volume) at a list of voxels
// (z, y, x) of the stored half-volume, in parallel over the list. A voxel outside the cutoff gets 0 and flag 0.
template<class Coord, class In, class Out, class W>
int run_insert_voxels(const noa_hip_project_args& a, const long long* zyx, long long n, void* values,
                      void* weights, signed char* inside, int n_threads) {
    const auto op = make_insert_op<Coord, In, Out, W>(a);
    In* vo = static_cast<In*>(values);
    W* wo = static_cast<W*>(weights);
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 16)
    for (long long i = 0; i < n; ++i) {
        In v = zero<In>();
        W w{};
        inside[i] = op.sum_at(zyx[3 * i], zyx[3 * i + 1], zyx[3 * i + 2], v, w) ? 1 : 0;
        vo[i] = v;
        wo[i] = w;
    }
    return 0;
}

template<class Coord, class In, class Out, class W>
int run_extract(const noa_hip_project_args& a, int n_threads, bool force) {
    ExtractOp<Coord, In, Out, W> op;
    op.volume = reader_volume<In>(a.input);
    op.has_wvolume = a.input_weight.data != nullptr;
    if (op.has_wvolume) op.wvolume = reader_volume<W>(a.input_weight);
    op.slices = writer_slices<Out>(a.output);
    op.has_wslices = a.output_weight.data != nullptr;
    if (op.has_wslices) op.wslices = writer_slices<W>(a.output_weight);
    op.geo = make_geometry<Coord>(a);
    for (int i = 0; i < 3; ++i) op.grid[i] = a.input_shape[i + 1];
    op.slice_h = a.output_shape[2];
    op.f_slice[0] = static_cast<Coord>(a.output_shape[2]);
    op.f_slice[1] = static_cast<Coord>(a.output_shape[3]);
    i64 t[3]; target3(a, a.input_shape, t);
    for (int i = 0; i < 3; ++i) op.f_target[i] = static_cast<Coord>(t[i]);
    op.cutoff_sqd = std::max(static_cast<Coord>(a.fftfreq_cutoff), Coord(0));
    op.cutoff_sqd *= op.cutoff_sqd;
    op.fftfreq_sinc = std::max(static_cast<Coord>(a.fftfreq_sinc), 1 / op.f_target[0]);
    op.fftfreq_blackman = std::max(static_cast<Coord>(a.fftfreq_blackman), 1 / op.f_target[0]);
    op.blackman_size = blackman_window_size<Coord>(op.fftfreq_blackman, op.f_target[0]);
    op.volume_centered = remap_src_centered(a.remap);
    op.slices_centered = remap_dst_centered(a.remap);
    op.interp = a.interp;
    const i64 B = a.output_shape[0], H = a.output_shape[2], Wr = a.output_shape[3] / 2 + 1;
    if (op.blackman_size > 1) {
        // ewise Zero of the outputs first (:1442-1446).
        for (i64 b = 0; b < B; ++b)
            for (i64 y = 0; y < H; ++y)
                for (i64 x = 0; x < Wr; ++x) {
                    op.slices(b, y, x) = zero<Out>();
                    if (op.has_wslices) op.wslices(b, y, x) = W{};
                }
        iwise4(B, op.blackman_size, H, Wr, op, n_threads, force);
    } else {
        iwise3(B, H, Wr, op, n_threads, force);
    }
    return 0;
}

// launch_insert_and_extract_central_slices_3d (:1489-1598) with the operator's constructor (:791-845).
template<class Coord, class In, class Out, class W>
int run_insert_extract(const noa_hip_insert_extract_args& a, int n_threads, bool force) {
    InsertExtractOp<Coord, In, Out, W> op;
    op.slices = reader_interp2<In>(a.input);
    op.has_wslices = a.input_weight.data != nullptr || a.input_weight.is_value;
    if (op.has_wslices) op.wslices = reader_interp2<W>(a.input_weight);
    op.out = writer_slices<Out>(a.output);
    op.has_wout = a.output_weight.data != nullptr;
    if (op.has_wout) op.wout = writer_slices<W>(a.output_weight);
    const bool general = has_ews(a.ews_radius) || has_scale(a.input_scaling) || has_scale(a.output_scaling);
    op.in_geo = make_geometry<Coord>(a.input_rotation, a.input_scaling, a.ews_radius, general);
    op.out_geo = make_geometry<Coord>(a.output_rotation, a.output_scaling, a.ews_radius, general);
    op.n_in = a.input_shape[0];
    op.in_shape[0] = a.input_shape[2];
    op.in_shape[1] = a.input_shape[3];
    op.f_in[0] = static_cast<Coord>(a.input_shape[2]);
    op.f_in[1] = static_cast<Coord>(a.input_shape[3]);
    op.out_h = a.output_shape[2];
    op.f_out[0] = static_cast<Coord>(a.output_shape[2]);
    op.f_out[1] = static_cast<Coord>(a.output_shape[3]);
    op.cutoff_sqd = std::max(static_cast<Coord>(a.fftfreq_cutoff), Coord(0));
    op.cutoff_sqd *= op.cutoff_sqd;
    op.volume_z = static_cast<Coord>(std::min(a.output_shape[2], a.output_shape[3]));   // no z: smallest axis
    op.in_sinc = std::max(static_cast<Coord>(a.input_fftfreq_sinc), 1 / op.volume_z);
    op.in_blackman = std::max(static_cast<Coord>(a.input_fftfreq_blackman), 1 / op.volume_z);
    op.w_sinc = std::max(static_cast<Coord>(a.w_fftfreq_sinc), 1 / op.volume_z);
    op.w_blackman = std::max(static_cast<Coord>(a.w_fftfreq_blackman), 1 / op.volume_z);
    op.w_size = blackman_window_size<Coord>(op.w_blackman, op.volume_z);
    op.in_centered = remap_src_centered(a.remap);
    op.out_centered = remap_dst_centered(a.remap);
    op.add = a.add_to_output != 0;
    op.correct = a.correct_weights != 0;
    op.interp = a.interp;
    const i64 B = a.output_shape[0], H = a.output_shape[2], Wr = a.output_shape[3] / 2 + 1;
    if (op.w_size > 1) {
        if (op.correct) { g_error = "oracle: correct_weights with a w window"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
        if (!op.add) {   // ewise Zero of the outputs first (:1543-1550)
            for (i64 b = 0; b < B; ++b)
                for (i64 y = 0; y < H; ++y)
                    for (i64 x = 0; x < Wr; ++x) {
                        op.out(b, y, x) = zero<Out>();
                        if (op.has_wout) op.wout(b, y, x) = W{};
                    }
        }
        iwise4(B, op.w_size, H, Wr, op, n_threads, force);
    } else {
        iwise3(B, H, Wr, op, n_threads, force);
    }
    return 0;
}

template<class Coord>
int dispatch_insert_extract(const noa_hip_insert_extract_args& a, int n_threads, bool force) {
    const int in = a.input.dtype, out = a.output.dtype;
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return run_insert_extract<Coord, Cx<float>, Cx<float>, float>(a, n_threads, force);
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return run_insert_extract<Coord, Cx<float>, float, float>(a, n_threads, force);
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return run_insert_extract<Coord, float, float, float>(a, n_threads, force);
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return run_insert_extract<Coord, Cx<double>, Cx<double>, double>(a, n_threads, force);
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return run_insert_extract<Coord, Cx<double>, double, double>(a, n_threads, force);
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return run_insert_extract<Coord, double, double, double>(a, n_threads, force);
    g_error = "oracle: unsupported value type combination";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

enum Op { RASTERIZE, INSERT, EXTRACT };

template<class Coord>
int dispatch_values(Op op, const noa_hip_project_args& a, int n_threads, bool force) {
    const int in = a.input.dtype, out = a.output.dtype;
    auto go = [&](auto in_tag, auto out_tag, auto w_tag) -> int {
        using In = decltype(in_tag);
        using Out = decltype(out_tag);
        using W = decltype(w_tag);
        switch (op) {
            case RASTERIZE: return run_rasterize<Coord, In, Out, W>(a, n_threads, force);
            case INSERT: return run_insert<Coord, In, Out, W>(a, n_threads, force);
            case EXTRACT: return run_extract<Coord, In, Out, W>(a, n_threads, force);
        }
        return 1;
    };
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return go(Cx<float>{}, Cx<float>{}, float{});
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return go(Cx<float>{}, float{}, float{});
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return go(float{}, float{}, float{});
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return go(Cx<double>{}, Cx<double>{}, double{});
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return go(Cx<double>{}, double{}, double{});
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return go(double{}, double{}, double{});
    g_error = "oracle: unsupported value type combination";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

inline bool weight_ok(const noa_hip_view& v, int real_dtype) {
    return (v.data == nullptr && !v.is_value) || v.dtype == real_dtype;
}

template<class Coord>
int dispatch_checked(Op op, const noa_hip_project_args& a, int n_threads, bool force) {
    const bool f64v = a.output.dtype == NOA_HIP_F64 || a.output.dtype == NOA_HIP_C64;
    const int wt = f64v ? NOA_HIP_F64 : NOA_HIP_F32;
    if (!weight_ok(a.input_weight, wt) || !weight_ok(a.output_weight, wt)) {
        g_error = "oracle: weights must be real with the precision of the values";
        return NOA_HIP_ERROR_UNSUPPORTED;
    }
    return dispatch_values<Coord>(op, a, n_threads, force);
    return NOA_HIP_ERROR_UNSUPPORTED;
}

int dispatch(Op op, const noa_hip_project_args* a, int n_threads, bool force) {
    g_error.clear();
    if (!a) { g_error = "oracle: null args"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const int base = a->interp >= 100 ? a->interp - 100 : a->interp;
    if (op != RASTERIZE && (base < NOA_HIP_INTERP_NEAREST || base > NOA_HIP_INTERP_LANCZOS8)) {
        g_error = "oracle: invalid interpolation method";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    const int k = a->rotation.kind;
    const bool f64 = k == NOA_HIP_MAT33_F64 || k == NOA_HIP_QUAT_F64;
    if (k < NOA_HIP_MAT33_F32 || k > NOA_HIP_QUAT_F64) { g_error = "oracle: invalid rotation kind"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    return f64 ? dispatch_checked<double>(op, *a, n_threads, force) : dispatch_checked<float>(op, *a, n_threads, force);
}

// ------------------------------------------------------------------------------------------------------------
// noa::signal::phase_shift_2d for rfft layouts (src/noa/signal/PhaseShift.hpp:78-139, :215-289), and
// nf::phase_shift (src/noa/fft/core/Frequency.hpp:228-248). Arrays: contiguous (B, H, W/2+1).
// ------------------------------------------------------------------------------------------------------------
template<class T>
constexpr T pi_v() { return static_cast<T>(3.141592653589793238462643383279502884); }

template<class V, class R>
int phase_shift_2d(const V* input, V* output, i64 b, i64 h, i64 w, int remap, const R* shifts, bool batched,
                   double cutoff) {
    using real_t = typename real_of<V>::type;
    const bool src_c = (remap & 4) != 0, dst_c = (remap & 8) != 0;
    const i64 wr = w / 2 + 1;
    // Single shift special cases (:250-267): zero -> copy/remap (or ones); |shift| == shape/2 and a cutoff that
    // covers the whole spectrum -> the +-1 pattern of PhaseShiftHalf (:18-69).
    int mode = 0;   // 0: general, 1: none, 2: half
    if (!batched) {
        auto allclose = [](R x, R y) {
            const R diff = std::abs(x - y);
            if (!std::isfinite(diff)) return false;
            return diff <= static_cast<R>(1e-6) || diff <= std::max(std::abs(x), std::abs(y)) * std::numeric_limits<R>::epsilon() * 2;
        };
        if (allclose(shifts[0], R(0)) && allclose(shifts[1], R(0))) mode = 1;
        else if (allclose(std::abs(shifts[0]), static_cast<R>(h) / 2) && allclose(std::abs(shifts[1]), static_cast<R>(w) / 2) &&
                 cutoff >= std::sqrt(0.5)) mode = 2;
    }
    const R norm_y = R(1) / static_cast<R>(h), norm_x = R(1) / static_cast<R>(w);   // m_norm = 1 / shape
    const R cutoff_sqd = static_cast<R>(cutoff) * static_cast<R>(cutoff);
    for (i64 bi = 0; bi < b; ++bi) {
        const R sy = shifts[batched ? 2 * bi : 0], sx = shifts[batched ? 2 * bi + 1 : 1];
        for (i64 y = 0; y < h; ++y) {
            const i64 fy = index2frequency(src_c, y, h);
            const i64 oy = frequency2index(dst_c, fy, h);
            for (i64 x = 0; x < wr; ++x) {
                V phase{};
                if constexpr (is_cx<V>::value) { phase.re = 1; phase.im = 0; }
                if (mode == 2) {
                    const real_t s = static_cast<real_t>((1 - 2 * std::abs(fy % 2)) * (1 - 2 * std::abs(x % 2)));
                    phase.re = s;
                    phase.im = 0;
                } else if (mode == 0) {
                    // fftfreq = frequency * (1 / shape), the z component being 0 (D == 1).
                    const R fz = R(0) * R(1), fyn = static_cast<R>(fy) * norm_y, fxn = static_cast<R>(x) * norm_x;
                    if ((fz * fz + fyn * fyn) + fxn * fxn <= cutoff_sqd) {
                        const R two_pi = 2 * pi_v<R>();
                        const R factor = -((two_pi * sy) * fyn + (two_pi * sx) * fxn);
                        const real_t f = static_cast<real_t>(factor);
                        phase.re = std::cos(f);
                        phase.im = std::sin(f);
                    }
                }
                V& o = output[(bi * h + oy) * wr + x];
                if (input) {
                    const V a = input[(bi * h + y) * wr + x];
                    o = V{a.re * phase.re - a.im * phase.im, a.re * phase.im + a.im * phase.re};
                } else {
                    o = phase;
                }
            }
        }
    }
    return NOA_HIP_OK;
}

// ------------------------------------------------------------------------------------------------------------
// noa::xform::fourier_interpolation_correction (src/noa/xform/FourierProject.hpp:1008-1061, :2062-2087): real-space
// correction of linear interpolation, x sinc^2(pi |r|) (pre) or / sinc^2 (post), r = (index - shape/2) / shape with
// half = 0 along dimensions of size 1. Contiguous (B, D, H, W) arrays, the input broadcast over the batch when
// its batch is 1.
// ------------------------------------------------------------------------------------------------------------
template<class In, class Out, class Coord>
void gridding_correction(const In* input, Out* output, i64 in_batch, const i64 shape[4], bool post) {
    const Coord fs[3] = {static_cast<Coord>(shape[1]), static_cast<Coord>(shape[2]), static_cast<Coord>(shape[3])};
    Coord half[3];
    for (int i = 0; i < 3; ++i) half[i] = fs[i] / 2 * static_cast<Coord>(shape[i + 1] != 1);
    const i64 n3 = shape[1] * shape[2] * shape[3];
    for (i64 b = 0; b < shape[0]; ++b)
        for (i64 j = 0; j < shape[1]; ++j)
            for (i64 k = 0; k < shape[2]; ++k)
                for (i64 l = 0; l < shape[3]; ++l) {
                    Coord d[3] = {static_cast<Coord>(j), static_cast<Coord>(k), static_cast<Coord>(l)};
                    for (int i = 0; i < 3; ++i) { d[i] -= half[i]; d[i] /= fs[i]; }
                    const Coord PI = static_cast<Coord>(3.141592653589793238462643383279502884);
                    const Coord radius = std::sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
                    const Coord x = PI * radius;
                    const Coord sinc = x == 0 ? Coord(1) : std::sin(x) / x;
                    const In sinc2 = static_cast<In>(sinc * sinc);
                    const i64 e = (j * shape[2] + k) * shape[3] + l;
                    const In value = input[(in_batch == 1 ? 0 : b) * n3 + e];
                    output[b * n3 + e] = static_cast<Out>(post ? value / sinc2 : value * sinc2);
                }
}

// nf::remap for the half-to-half rfft layouts (fft/Remap.hpp, fft/core/Frequency.hpp:198-225): out[b, remap(z, y),
// x] = in[b, z, y, x], the z and y indices fftshifted (non-centred -> centred) or ifftshifted. Contiguous arrays.
template<class T>
void remap_hx2hx(const T* input, T* output, const i64 shape[4], int remap) {
    const bool src_c = (remap & 4) != 0, dst_c = (remap & 8) != 0;
    const i64 wr = shape[3] / 2 + 1;
    for (i64 b = 0; b < shape[0]; ++b)
        for (i64 z = 0; z < shape[1]; ++z) {
            const i64 oz = frequency2index(dst_c, index2frequency(src_c, z, shape[1]), shape[1]);
            for (i64 y = 0; y < shape[2]; ++y) {
                const i64 oy = frequency2index(dst_c, index2frequency(src_c, y, shape[2]), shape[2]);
                for (i64 x = 0; x < wr; ++x)
                    output[((b * shape[1] + oz) * shape[2] + oy) * wr + x] = input[((b * shape[1] + z) * shape[2] + y) * wr + x];
            }
        }
}

// ------------------------------------------------------------------------------------------------------------
// noa::signal::CTFAnisotropic<T> (src/noa/signal/core/CTF.hpp:291-575) and the CTF operator for 2D rfft layouts
// (src/noa/signal/CTF.hpp:14-125, public function :430-500).
// ------------------------------------------------------------------------------------------------------------
template<class T>
struct alignas(2 * sizeof(T)) CtfAniso {   // members in declaration order (:564-575)
    T pixel_size[2];
    T defocus_value, defocus_astigmatism, defocus_angle;
    T phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
};

// relativistic_electron_wavelength (core/CTF.hpp:14-25), constants of base/Math.hpp:13-17.
inline double relativistic_electron_wavelength(double energy) {
    const double h = 6.62607015e-34, c = 299792458.0, m0 = 9.1093837015e-31, e = 1.602176634e-19;
    const double eV = e * energy;
    const double numerator = h * c;
    const double denominator = std::sqrt(eV * (2 * m0 * c * c + eV));
    return numerator / denominator;
}

// The constructor from physical parameters (:318-345): pixel size (A/px), defocus (um, positive underfocus),
// astigmatism (um), its angle (rad), voltage (kV), amplitude fraction, cs (mm), phase shift (rad), bfactor (A^2),
// scale.
template<class T>
CtfAniso<T> make_ctf(const T pixel_size[2], T defocus, T astigmatism, T angle, T voltage, T amplitude, T cs,
                     T phase_shift, T bfactor, T scale) {
    CtfAniso<T> c{};
    c.pixel_size[0] = pixel_size[0];
    c.pixel_size[1] = pixel_size[1];
    c.phase_shift = phase_shift;
    c.cs = cs * static_cast<T>(1e7);
    c.voltage = voltage * static_cast<T>(1e3);
    c.amplitude = amplitude;
    c.scale = scale;
    c.defocus_value = -defocus * static_cast<T>(1e4);         // set_defocus (:405-411)
    c.defocus_astigmatism = -astigmatism * static_cast<T>(1e4);
    c.defocus_angle = angle;
    c.bfactor_forth = bfactor / 4;                             // set_bfactor (:401-404)
    const double lambda = relativistic_electron_wavelength(static_cast<double>(c.voltage)) * 1e10;   // (:545-551)
    const double PI = 3.1415926535897932384626433832795;
    c.k1 = static_cast<T>(PI * lambda);
    c.k2 = static_cast<T>(PI * 0.5 * static_cast<double>(c.cs) * lambda * lambda * lambda);
    c.k3 = std::atan(c.amplitude / std::sqrt(1 - c.amplitude * c.amplitude));   // (:553-555)
    return c;
}

// value_at (:473-488) with phi2defocus_ (:540-543), cartesian2phi<false> / cartesian2rho (xform/core/Polar.hpp).
template<class T>
T ctf_value_at(const CtfAniso<T>& c, T fy, T fx) {
    const T sy = fy / c.pixel_size[0], sx = fx / c.pixel_size[1];
    const T phi = std::atan2(sy, sx);
    const T rho = std::sqrt(sy * sy + sx * sx);
    const T r1 = rho;
    const T r2 = r1 * r1;
    const T r4 = r2 * r2;
    const T defocus = c.defocus_value + c.defocus_astigmatism * std::cos(2 * (phi - c.defocus_angle));
    const T phase = c.k1 * r2 * defocus + c.k2 * r4 - c.phase_shift - c.k3;
    T ctf = -std::sin(phase);
    if (c.bfactor_forth != 0) ctf *= std::exp(c.bfactor_forth * r2);
    return ctf * c.scale;
}

// The operator on contiguous (b, h, w/2+1) arrays; `input` may be null (the CTF is written), and has a batch of
// `in_batch` (1: broadcast). Linspace::for_size (runtime/core/Iwise.hpp:67-78) gives the fftfreq step per axis.
template<class T, class In, class Out, bool HAS_INPUT>
void ctf_anisotropic_2d(const In* input, i64 in_batch, Out* output, i64 b, i64 h, i64 w, int remap,
                        const CtfAniso<T>* ctfs, bool batched, double start, double stop, bool endpoint,
                        bool ctf_abs, bool ctf_squared) {
    const bool src_c = (remap & 4) != 0, dst_c = (remap & 8) != 0;
    const i64 wr = w / 2 + 1;
    T step[2];
    const i64 dims[2] = {h, w};
    for (int i = 0; i < 2; ++i) {
        const i64 size = dims[i] / 2 + 1;
        const T t_start = static_cast<T>(start), t_stop_opt = static_cast<T>(stop);
        const T t_stop = t_stop_opt <= 0 ? static_cast<T>(dims[i] / 2) / static_cast<T>(dims[i]) : t_stop_opt;
        const i64 count = std::max<i64>(1, size - (endpoint ? 1 : 0));
        step[i] = (t_stop - t_start) / static_cast<T>(count);
    }
    using RIn = typename real_of<In>::type;
    for (i64 bi = 0; bi < b; ++bi) {
        const CtfAniso<T>& c = ctfs[batched ? bi : 0];
        for (i64 y = 0; y < h; ++y) {
            const i64 fy = index2frequency(dst_c, y, h);
            const i64 iy = frequency2index(src_c, fy, h);
            for (i64 x = 0; x < wr; ++x) {
                T ctf = ctf_value_at(c, static_cast<T>(fy) * step[0], static_cast<T>(x) * step[1]);
                if (ctf_squared) ctf *= ctf;
                else if (ctf_abs) ctf = std::abs(ctf);
                Out& o = output[(bi * h + y) * wr + x];
                if constexpr (HAS_INPUT) {
                    o = cast_or_abs_squared<Out>(input[((in_batch == 1 ? 0 : bi) * h + iy) * wr + x] * static_cast<RIn>(ctf));
                } else {
                    if constexpr (is_cx<Out>::value) o = Out{static_cast<typename real_of<Out>::type>(ctf), 0};
                    else o = static_cast<Out>(ctf);
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------------------------
// Cubic B-spline prefilter (xform/CubicBSplinePrefilter.hpp:24, cpu/CubicBSplinePrefilter.hpp:9-158, 1D recursion
// xform/core/CubicBSplinePrefilter.hpp:22-107): the causal then anticausal recursion of the cubic B-spline pole
// z = sqrt(3) - 2 along each line, gain (1 - z)(1 - 1/z), the causal initialisation summing the first
// min(12, n) samples (half-sample symmetric boundary), the anticausal one z / (z - 1) * c[n-1]. Lines along x,
// then y, then z, after the dimensions are reordered to the output's rightmost order; ndim() of the reordered
// shape picks 1, 2 or 3 passes. Element strides; input == output filters in place.
// ------------------------------------------------------------------------------------------------------------
template<class T>
struct BSpline1d {
    using R = typename real_of<T>::type;
    static constexpr R POLE = static_cast<R>(-0.2679491924311228);
    static constexpr R LAMBDA = (1 - POLE) * (1 - 1 / POLE);

    static T mul(R s, T v) {
        if constexpr (is_cx<T>::value) return T{s * v.re, s * v.im};
        else return s * v;
    }
    static T sub(T a, T b) {
        if constexpr (is_cx<T>::value) return T{a.re - b.re, a.im - b.im};
        else return a - b;
    }
    // Causal initialisation (core/CubicBSplinePrefilter.hpp:87-100).
    static T initial_causal(const T* c, i64 stride, i64 size) {
        const i64 horizon = std::min<i64>(12, size);
        R zn = POLE;
        T sum = *c;
        for (i64 n = 0; n < horizon; ++n) {
            sum = sum + mul(zn, *c);
            zn *= POLE;
            c += stride;
        }
        return sum;
    }
    // filter() (:60-80); filter_inplace() is the same with input == output.
    static void filter(const T* input, i64 is, T* output, i64 os, i64 size) {
        T* c = output;
        T prev;
        *c = prev = mul(LAMBDA, initial_causal(input, is, size));
        for (i64 n = 1; n < size; ++n) {
            input += is;
            c += os;
            *c = prev = mul(LAMBDA, *input) + mul(POLE, prev);
        }
        *c = prev = mul(POLE / (POLE - 1), *c);
        for (i64 n = size - 2; 0 <= n; --n) {
            c -= os;
            *c = prev = mul(POLE, sub(prev, *c));
        }
    }
};

// Shape::ndim() (runtime/core/Shape.hpp:315-335) of a BDHW shape.
inline int ndim4(const i64 s[4]) {
    if (s[1] > 1) return 3;
    if (s[2] > 1 && s[3] > 1) return 2;
    return 1;
}

template<class T>
void cubic_bspline_prefilter(const T* input, const i64 is_[4], T* output, const i64 os_[4], const i64 shape_[4]) {
    using F = BSpline1d<T>;
    i64 is[4], os[4], sh[4];
    for (int i = 0; i < 4; ++i) { is[i] = is_[i]; os[i] = os_[i]; sh[i] = shape_[i]; }
    // Strides::rightmost_order (Shape.hpp:871-886): stable sort of D, H, W by decreasing output stride, a dimension
    // of size 1 counting as the largest stride.
    int order[3] = {0, 1, 2};
    i64 key[3];
    for (int i = 0; i < 3; ++i) key[i] = sh[i + 1] <= 1 ? std::numeric_limits<i64>::max() : os[i + 1];
    std::stable_sort(order, order + 3, [&](int a, int b) { return key[a] > key[b]; });
    {
        i64 ti[4] = {is[0]}, to[4] = {os[0]}, ts[4] = {sh[0]};
        for (int i = 0; i < 3; ++i) { ti[i + 1] = is[order[i] + 1]; to[i + 1] = os[order[i] + 1]; ts[i + 1] = sh[order[i] + 1]; }
        for (int i = 0; i < 4; ++i) { is[i] = ti[i]; os[i] = to[i]; sh[i] = ts[i]; }
    }
    const bool inplace = static_cast<const void*>(input) == static_cast<const void*>(output);
    // Rows (x), from the input (or in place).
    auto rows = [&](i64 n_outer0, i64 n_outer1, i64 n_outer2, const i64 ib[3], const i64 ob[3], i64 isx, i64 osx,
                    i64 n) {
        for (i64 a = 0; a < n_outer0; ++a)
            for (i64 b = 0; b < n_outer1; ++b)
                for (i64 c = 0; c < n_outer2; ++c) {
                    T* o = output + a * ob[0] + b * ob[1] + c * ob[2];
                    if (inplace) F::filter(o, osx, o, osx, n);
                    else F::filter(input + a * ib[0] + b * ib[1] + c * ib[2], isx, o, osx, n);
                }
    };
    // In place along one axis of the output.
    auto lines = [&](i64 n0, i64 n1, i64 n2, i64 s0, i64 s1, i64 s2, i64 stride, i64 n) {
        for (i64 a = 0; a < n0; ++a)
            for (i64 b = 0; b < n1; ++b)
                for (i64 c = 0; c < n2; ++c) {
                    T* o = output + a * s0 + b * s1 + c * s2;
                    F::filter(o, stride, o, stride, n);
                }
    };
    const int nd = ndim4(sh);
    if (nd == 3) {
        const i64 ib[3] = {is[0], is[1], is[2]}, ob[3] = {os[0], os[1], os[2]};
        rows(sh[0], sh[1], sh[2], ib, ob, is[3], os[3], sh[3]);                 // every row
        lines(sh[0], sh[1], sh[3], os[0], os[1], os[3], os[2], sh[2]);          // every column
        lines(sh[0], sh[2], sh[3], os[0], os[2], os[3], os[1], sh[1]);          // every page
    } else if (nd == 2) {
        const i64 ib[3] = {is[0], is[2], 0}, ob[3] = {os[0], os[2], 0};
        rows(sh[0], sh[2], 1, ib, ob, is[3], os[3], sh[3]);
        lines(sh[0], sh[3], 1, os[0], os[3], 0, os[2], sh[2]);
    } else {
        const i64 ib[3] = {is[0], 0, 0}, ob[3] = {os[0], 0, 0};
        rows(sh[0], 1, 1, ib, ob, is[3], os[3], sh[3]);
    }
}

// ------------------------------------------------------------------------------------------------------------
// noa::signal::CTFIsotropic<T> (src/noa/signal/core/CTF.hpp:30-272) and the CTF operator over 1, 2 or 3
// dimensions for every half and full layout (src/noa/signal/CTF.hpp:14-125), with the dispatch of the public
// functions (ctf_isotropic :216-421, ctf_anisotropic :430-500): the dimensions are put in the rightmost order of
// the output first, the input's (broadcast) strides following.
// ------------------------------------------------------------------------------------------------------------
template<class T>
struct CtfIso {   // members in declaration order (:259-271)
    T pixel_size, defocus, phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
};

// The constructor from physical parameters (:61-82): pixel size (A/px), defocus (um, positive underfocus),
// voltage (kV), amplitude fraction, cs (mm), phase shift (rad), bfactor (A^2), scale.
template<class T>
CtfIso<T> make_ctf_iso(T pixel_size, T defocus, T voltage, T amplitude, T cs, T phase_shift, T bfactor, T scale) {
    CtfIso<T> c{};
    c.pixel_size = pixel_size;
    c.defocus = -defocus * static_cast<T>(1e4);
    c.phase_shift = phase_shift;
    c.cs = cs * static_cast<T>(1e7);
    c.voltage = voltage * static_cast<T>(1e3);
    c.amplitude = amplitude;
    c.scale = scale;
    c.bfactor_forth = bfactor / 4;                               // set_bfactor (:147-150)
    const double lambda = relativistic_electron_wavelength(static_cast<double>(c.voltage)) * 1e10;   // (:246-252)
    const double PI = 3.1415926535897932384626433832795;
    c.k1 = static_cast<T>(PI * lambda);
    c.k2 = static_cast<T>(PI * 0.5 * static_cast<double>(c.cs) * lambda * lambda * lambda);
    c.k3 = std::atan(c.amplitude / std::sqrt(1 - c.amplitude * c.amplitude));   // (:254-256)
    return c;
}

// value_at (:204-212).
template<class T>
T ctf_value_at(const CtfIso<T>& c, T fftfreq) {
    const T r1 = fftfreq / c.pixel_size;
    const T r2 = r1 * r1;
    const T r4 = r2 * r2;
    const T phase = (c.k1 * r2 * c.defocus) + (c.k2 * r4) - c.phase_shift - c.k3;
    T ctf = -std::sin(phase);
    if (c.bfactor_forth != 0) ctf *= std::exp(c.bfactor_forth * r2);
    return ctf * c.scale;
}

// Linspace{start, stop <= 0 ? highest_fftfreq(n) : stop, endpoint}.for_size(n / 2 + 1).step (:65-77).
template<class T>
T ctf_fftfreq_step(i64 n, double start, double stop, bool endpoint) {
    const i64 size = n / 2 + 1;
    const T t_start = static_cast<T>(start), t_stop_opt = static_cast<T>(stop);
    const T t_stop = t_stop_opt <= 0 ? static_cast<T>(n / 2) / static_cast<T>(n) : t_stop_opt;
    const i64 count = std::max<i64>(1, size - (endpoint ? 1 : 0));
    return (t_stop - t_start) / static_cast<T>(count);
}

// Returns 0, or 1 when the reordered iteration would leave the output (noa itself would write out of bounds: an
// rfft output whose half dimension is not the innermost one).
template<class T, class In, class Out, class Ctf, bool HAS_INPUT>
int ctf_general(const In* input, const i64 is_[4], Out* output, const i64 os_[4], const i64 out_shape[4],
                const i64 shape_[4], int remap, const Ctf* ctfs, bool batched, double start, double stop,
                bool endpoint, bool ctf_abs, bool ctf_squared) {
    constexpr bool ISO = std::is_same<Ctf, CtfIso<T>>::value;
    const bool rfft = (remap & 2) == 0, src_c = (remap & 4) != 0, dst_c = (remap & 8) != 0;
    i64 is[4], os[4], sh[4], ext[4];   // ext: the output's extent along each (reordered) dimension
    for (int i = 0; i < 4; ++i) { is[i] = is_[i]; os[i] = os_[i]; sh[i] = shape_[i]; ext[i] = out_shape[i]; }
    const int nd = ndim4(sh);
    int N = nd;
    i64 dims[3]{}, ist[3]{}, ost[3]{}, dext[3]{};
    if (nd == 1) {                       // non_empty_dhw_dimension (runtime/core/Utils.hpp:30-35)
        int index = 3;
        for (int i = 1; i < 4; ++i) if (sh[i] > 1) { index = i; break; }
        dims[0] = sh[index]; ist[0] = is[index]; ost[0] = os[index]; dext[0] = ext[index];
    } else if (nd == 2) {                // H and W swapped unless rightmost (:270-276)
        const i64 k2 = sh[2] <= 1 ? std::numeric_limits<i64>::max() : os[2];
        const i64 k3 = sh[3] <= 1 ? std::numeric_limits<i64>::max() : os[3];
        if (!(k2 >= k3)) { std::swap(is[2], is[3]); std::swap(os[2], os[3]); std::swap(sh[2], sh[3]); std::swap(ext[2], ext[3]); }
        for (int i = 0; i < 2; ++i) { dims[i] = sh[i + 2]; ist[i] = is[i + 2]; ost[i] = os[i + 2]; dext[i] = ext[i + 2]; }
    } else {                             // permute_all to the rightmost order (:298-303)
        int order[3] = {0, 1, 2};
        i64 key[3];
        for (int i = 0; i < 3; ++i) key[i] = sh[i + 1] <= 1 ? std::numeric_limits<i64>::max() : os[i + 1];
        std::stable_sort(order, order + 3, [&](int a, int b) { return key[a] > key[b]; });
        for (int i = 0; i < 3; ++i) {
            dims[i] = sh[order[i] + 1]; ist[i] = is[order[i] + 1]; ost[i] = os[order[i] + 1]; dext[i] = ext[order[i] + 1];
        }
    }
    T step[3];
    for (int i = 0; i < N; ++i) step[i] = ctf_fftfreq_step<T>(dims[i], start, stop, endpoint);
    i64 iter[3];
    for (int i = 0; i < N; ++i) iter[i] = (rfft && i == N - 1) ? dims[i] / 2 + 1 : dims[i];
    for (int i = 0; i < N; ++i) if (iter[i] > dext[i]) return 1;
    const T t_start = static_cast<T>(start);
    using RIn = typename real_of<In>::type;
    i64 idx[3] = {0, 0, 0};
    const i64 total = iter[0] * (N > 1 ? iter[1] : 1) * (N > 2 ? iter[2] : 1);
    for (i64 b = 0; b < sh[0]; ++b) {
        const Ctf& c = ctfs[batched ? b : 0];
        for (i64 t = 0; t < total; ++t) {
            i64 r = t;
            for (int i = N - 1; i >= 0; --i) { idx[i] = r % iter[i]; r /= iter[i]; }
            T f[3];
            i64 in_off = b * is[0], out_off = b * os[0];
            for (int i = 0; i < N; ++i) {
                const bool half = rfft && i == N - 1;
                const i64 freq = half ? idx[i] : index2frequency(dst_c, idx[i], dims[i]);
                const i64 src = half ? idx[i] : frequency2index(src_c, freq, dims[i]);   // remap_indices
                f[i] = static_cast<T>(freq) * step[i];
                in_off += src * ist[i];
                out_off += idx[i] * ost[i];
            }
            T ctf;
            if constexpr (ISO) {
                T arg;
                if (N == 1) arg = f[0] + t_start;
                else if (N == 2) arg = std::sqrt(f[0] * f[0] + f[1] * f[1]);
                else arg = std::sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
                ctf = ctf_value_at(c, arg);
            } else {
                ctf = ctf_value_at(c, f[0], f[1]);
            }
            if (ctf_squared) ctf *= ctf;
            else if (ctf_abs) ctf = std::abs(ctf);
            Out& o = output[out_off];
            if constexpr (HAS_INPUT) {
                o = cast_or_abs_squared<Out>(input[in_off] * static_cast<RIn>(ctf));
            } else {
                if constexpr (is_cx<Out>::value) o = Out{static_cast<typename real_of<Out>::type>(ctf), 0};
                else o = static_cast<Out>(ctf);
            }
        }
    }
    return 0;
}

} // namespace oracle

extern "C" {

/* noa::xform::cubic_bspline_prefilter on strided (B, D, H, W) arrays; dtype 1 f32, 2 f64, 3 c32, 4 c64. Strides in
 * elements; input == output filters in place. */
int oracle_cubic_bspline_prefilter(const void* input, const long long in_strides[4], void* output,
                                   const long long out_strides[4], const long long shape[4], int dtype) {
    using namespace oracle;
    i64 is[4], os[4], sh[4];
    for (int i = 0; i < 4; ++i) { is[i] = in_strides[i]; os[i] = out_strides[i]; sh[i] = shape[i]; }
    switch (dtype) {
        case 1: cubic_bspline_prefilter(static_cast<const float*>(input), is, static_cast<float*>(output), os, sh); return 0;
        case 2: cubic_bspline_prefilter(static_cast<const double*>(input), is, static_cast<double*>(output), os, sh); return 0;
        case 3: cubic_bspline_prefilter(static_cast<const Cx<float>*>(input), is, static_cast<Cx<float>*>(output), os, sh); return 0;
        case 4: cubic_bspline_prefilter(static_cast<const Cx<double>*>(input), is, static_cast<Cx<double>*>(output), os, sh); return 0;
        default: return NOA_HIP_ERROR_UNSUPPORTED;
    }
}

const char* oracle_last_error(void) { return oracle::g_error.c_str(); }

/* n_threads: thread limit (noa's Session thread limit); force != 0 uses exactly n_threads (ignores the 2^20
 * per-thread heuristic of cpu/Iwise.hpp:104-116). */
int oracle_rasterize_central_slices_3d(const noa_hip_project_args* a, int n_threads, int force) {
    return oracle::dispatch(oracle::RASTERIZE, a, n_threads, force != 0);
}
int oracle_insert_central_slices_3d(const noa_hip_project_args* a, int n_threads, int force) {
    return oracle::dispatch(oracle::INSERT, a, n_threads, force != 0);
}
// Per-voxel sums of the interpolating insertion at a list of (z, y, x) voxels (values: the input's value type,
// weights: its real type); `inside` receives 1 for voxels inside the cutoff. Single-precision or double rotations.
int oracle_insert_central_slices_3d_voxels(const noa_hip_project_args* a, const long long* zyx, long long n,
                                           void* values, void* weights, signed char* inside, int n_threads) {
    using namespace oracle;
    g_error.clear();
    const bool f64 = a->rotation.kind == NOA_HIP_MAT33_F64 || a->rotation.kind == NOA_HIP_QUAT_F64;
    const int in = a->input.dtype;
    auto go = [&](auto coord) -> int {
        using Coord = decltype(coord);
        if (in == NOA_HIP_C32) return run_insert_voxels<Coord, Cx<float>, Cx<float>, float>(*a, zyx, n, values, weights, inside, n_threads);
        if (in == NOA_HIP_F32) return run_insert_voxels<Coord, float, float, float>(*a, zyx, n, values, weights, inside, n_threads);
        if (in == NOA_HIP_C64) return run_insert_voxels<Coord, Cx<double>, Cx<double>, double>(*a, zyx, n, values, weights, inside, n_threads);
        if (in == NOA_HIP_F64) return run_insert_voxels<Coord, double, double, double>(*a, zyx, n, values, weights, inside, n_threads);
        g_error = "oracle: unsupported value type";
        return NOA_HIP_ERROR_UNSUPPORTED;
    };
    return f64 ? go(double{}) : go(float{});
}
int oracle_extract_central_slices_3d(const noa_hip_project_args* a, int n_threads, int force) {
    return oracle::dispatch(oracle::EXTRACT, a, n_threads, force != 0);
}

/* insert_and_extract_central_slices_3d; the coordinate precision is the input rotation's (:1514). */
int oracle_insert_and_extract_central_slices_3d(const noa_hip_insert_extract_args* a, int n_threads, int force) {
    using namespace oracle;
    g_error.clear();
    if (!a) { g_error = "oracle: null args"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const int base = a->interp >= 100 ? a->interp - 100 : a->interp;
    if (base < NOA_HIP_INTERP_NEAREST || base > NOA_HIP_INTERP_LANCZOS8) {
        g_error = "oracle: invalid interpolation method";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    const int k = a->input_rotation.kind;
    if (k < NOA_HIP_MAT33_F32 || k > NOA_HIP_QUAT_F64) { g_error = "oracle: invalid rotation kind"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    const bool f64 = k == NOA_HIP_MAT33_F64 || k == NOA_HIP_QUAT_F64;
    return f64 ? dispatch_insert_extract<double>(*a, n_threads, force != 0)
               : dispatch_insert_extract<float>(*a, n_threads, force != 0);
}

/* noa::signal::phase_shift_2d<REMAP> on contiguous (b, h, w/2+1) c32 or c64 arrays; `input` may be NULL (the
 * phase shifts are written); shifts: (y, x) pairs, f32 or f64, one (batched == 0) or b of them. */
int oracle_phase_shift_2d(const void* input, void* output, int c64, long long b, long long h, long long w, int remap,
                          const void* shifts, int shifts_f64, int batched, double cutoff) {
    using namespace oracle;
    if (c64) {
        if (shifts_f64) return phase_shift_2d(static_cast<const Cx<double>*>(input), static_cast<Cx<double>*>(output), b, h, w, remap, static_cast<const double*>(shifts), batched != 0, cutoff);
        return phase_shift_2d(static_cast<const Cx<double>*>(input), static_cast<Cx<double>*>(output), b, h, w, remap, static_cast<const float*>(shifts), batched != 0, cutoff);
    }
    if (shifts_f64) return phase_shift_2d(static_cast<const Cx<float>*>(input), static_cast<Cx<float>*>(output), b, h, w, remap, static_cast<const double*>(shifts), batched != 0, cutoff);
    return phase_shift_2d(static_cast<const Cx<float>*>(input), static_cast<Cx<float>*>(output), b, h, w, remap, static_cast<const float*>(shifts), batched != 0, cutoff);
}

/* fourier_interpolation_correction on contiguous (B, D, H, W) real arrays; f32 or f64 in and out (FP64 math if
 * either is f64). */
int oracle_fourier_interpolation_correction(const void* input, int in_f64, long long in_batch, void* output,
                                            int out_f64, const long long shape[4], int post) {
    using namespace oracle;
    const i64 sh[4] = {shape[0], shape[1], shape[2], shape[3]};
    if (in_f64 && out_f64) gridding_correction<double, double, double>(static_cast<const double*>(input), static_cast<double*>(output), in_batch, sh, post != 0);
    else if (in_f64) gridding_correction<double, float, double>(static_cast<const double*>(input), static_cast<float*>(output), in_batch, sh, post != 0);
    else if (out_f64) gridding_correction<float, double, double>(static_cast<const float*>(input), static_cast<double*>(output), in_batch, sh, post != 0);
    else gridding_correction<float, float, float>(static_cast<const float*>(input), static_cast<float*>(output), in_batch, sh, post != 0);
    return 0;
}

/* nf::remap between half-to-half layouts on contiguous rfft arrays of `elem_bytes` 4, 8 or 16. */
int oracle_remap_hx2hx(const void* input, void* output, int elem_bytes, const long long shape[4], int remap) {
    using namespace oracle;
    const i64 sh[4] = {shape[0], shape[1], shape[2], shape[3]};
    if (elem_bytes == 4) remap_hx2hx(static_cast<const float*>(input), static_cast<float*>(output), sh, remap);
    else if (elem_bytes == 8) remap_hx2hx(static_cast<const double*>(input), static_cast<double*>(output), sh, remap);
    else if (elem_bytes == 16) remap_hx2hx(static_cast<const Cx<double>*>(input), static_cast<Cx<double>*>(output), sh, remap);
    else return NOA_HIP_ERROR_UNSUPPORTED;
    return 0;
}

/* noa::signal::ctf_anisotropic<REMAP> on contiguous (b, h, w/2+1) arrays. dtypes: 1 f32, 2 f64, 3 c32, 4 c64
 * (input 0: none); ctf_f64 selects CTFAnisotropic<f64>; `ctfs` holds one (batched == 0) or b objects. */
int oracle_ctf_anisotropic_2d(const void* input, int in_dtype, long long in_batch, void* output, int out_dtype,
                              long long b, long long h, long long w, int remap, const void* ctfs, int ctf_f64,
                              int batched, double start, double stop, int endpoint, int ctf_abs, int ctf_squared) {
    using namespace oracle;
    auto go = [&](auto t_tag) -> int {
        using T = decltype(t_tag);
        const auto* c = static_cast<const CtfAniso<T>*>(ctfs);
        auto run = [&](auto in_tag, auto out_tag) {
            using In = decltype(in_tag);
            using Out = decltype(out_tag);
            if (in_dtype == 0)
                ctf_anisotropic_2d<T, In, Out, false>(nullptr, in_batch, static_cast<Out*>(output), b, h, w, remap, c,
                                                      batched != 0, start, stop, endpoint != 0, ctf_abs != 0,
                                                      ctf_squared != 0);
            else if constexpr (is_cx<In>::value || !is_cx<Out>::value)
                ctf_anisotropic_2d<T, In, Out, true>(static_cast<const In*>(input), in_batch, static_cast<Out*>(output), b, h, w,
                                                     remap, c, batched != 0, start, stop, endpoint != 0, ctf_abs != 0,
                                                     ctf_squared != 0);
            return 0;
        };
        if (in_dtype == 3 && out_dtype == 3) return run(Cx<float>{}, Cx<float>{});
        if (in_dtype == 3 && out_dtype == 1) return run(Cx<float>{}, float{});
        if (in_dtype == 1 && out_dtype == 1) return run(float{}, float{});
        if (in_dtype == 4 && out_dtype == 4) return run(Cx<double>{}, Cx<double>{});
        if (in_dtype == 4 && out_dtype == 2) return run(Cx<double>{}, double{});
        if (in_dtype == 2 && out_dtype == 2) return run(double{}, double{});
        if (in_dtype == 0 && out_dtype == 1) return run(float{}, float{});
        if (in_dtype == 0 && out_dtype == 2) return run(double{}, double{});
        if (in_dtype == 0 && out_dtype == 3) return run(float{}, Cx<float>{});
        if (in_dtype == 0 && out_dtype == 4) return run(double{}, Cx<double>{});
        g_error = "oracle: unsupported CTF value types";
        return NOA_HIP_ERROR_UNSUPPORTED;
    };
    return ctf_f64 ? go(double{}) : go(float{});
}

/* CTFAnisotropic<T> from physical parameters, written to `out` (56 or 112 bytes). */
void oracle_ctf_anisotropic_make(int f64, const double pixel_size[2], double defocus, double astigmatism,
                                 double angle, double voltage, double amplitude, double cs, double phase_shift,
                                 double bfactor, double scale, void* out) {
    using namespace oracle;
    if (f64) {
        const double ps[2] = {pixel_size[0], pixel_size[1]};
        const auto c = make_ctf<double>(ps, defocus, astigmatism, angle, voltage, amplitude, cs, phase_shift, bfactor, scale);
        std::memcpy(out, &c, sizeof(c));
    } else {
        const float ps[2] = {static_cast<float>(pixel_size[0]), static_cast<float>(pixel_size[1])};
        const auto c = make_ctf<float>(ps, static_cast<float>(defocus), static_cast<float>(astigmatism),
                                       static_cast<float>(angle), static_cast<float>(voltage),
                                       static_cast<float>(amplitude), static_cast<float>(cs),
                                       static_cast<float>(phase_shift), static_cast<float>(bfactor),
                                       static_cast<float>(scale));
        std::memcpy(out, &c, sizeof(c));
    }
}
double oracle_relativistic_electron_wavelength(double volts) { return oracle::relativistic_electron_wavelength(volts); }

/* CTFIsotropic<T> from physical parameters, written to `out` (44 or 88 bytes). */
void oracle_ctf_isotropic_make(int f64, double pixel_size, double defocus, double voltage, double amplitude, double cs,
                               double phase_shift, double bfactor, double scale, void* out) {
    using namespace oracle;
    if (f64) {
        const auto c = make_ctf_iso<double>(pixel_size, defocus, voltage, amplitude, cs, phase_shift, bfactor, scale);
        std::memcpy(out, &c, sizeof(c));
    } else {
        const auto c = make_ctf_iso<float>(static_cast<float>(pixel_size), static_cast<float>(defocus),
                                           static_cast<float>(voltage), static_cast<float>(amplitude),
                                           static_cast<float>(cs), static_cast<float>(phase_shift),
                                           static_cast<float>(bfactor), static_cast<float>(scale));
        std::memcpy(out, &c, sizeof(c));
    }
}

/* noa::signal::ctf_isotropic (isotropic != 0: CTFIsotropic<T> objects over 1, 2 or 3 dimensions) or ctf_anisotropic
 * (CTFAnisotropic<T>, 2D) on strided arrays, every half and full layout. shape: the logical BDHW shape; out_shape:
 * the output's shape (rfft or full); strides in elements (the input's already broadcast to the output). dtypes: 1 f32,
 * 2 f64, 3 c32, 4 c64 (input 0: none). Returns 1 if the reordered iteration would leave the output. */
int oracle_ctf(int isotropic, const void* input, int in_dtype, const long long in_strides[4], void* output,
               int out_dtype, const long long out_strides[4], const long long out_shape[4], const long long shape[4],
               int remap, const void* ctfs, int ctf_f64, int batched, double start, double stop, int endpoint,
               int ctf_abs, int ctf_squared) {
    using namespace oracle;
    i64 is[4], os[4], osh[4], sh[4];
    for (int i = 0; i < 4; ++i) { is[i] = in_strides[i]; os[i] = out_strides[i]; osh[i] = out_shape[i]; sh[i] = shape[i]; }
    auto go = [&](auto t_tag, auto ctf_tag) -> int {
        using T = decltype(t_tag);
        using Ctf = decltype(ctf_tag);
        const auto* c = static_cast<const Ctf*>(ctfs);
        auto run = [&](auto in_tag, auto out_tag) -> int {
            using In = decltype(in_tag);
            using Out = decltype(out_tag);
            if (in_dtype == 0)
                return ctf_general<T, In, Out, Ctf, false>(nullptr, is, static_cast<Out*>(output), os, osh, sh, remap, c,
                                                           batched != 0, start, stop, endpoint != 0, ctf_abs != 0,
                                                           ctf_squared != 0);
            if constexpr (is_cx<In>::value || !is_cx<Out>::value)
                return ctf_general<T, In, Out, Ctf, true>(static_cast<const In*>(input), is, static_cast<Out*>(output), os,
                                                          osh, sh, remap, c, batched != 0, start, stop, endpoint != 0,
                                                          ctf_abs != 0, ctf_squared != 0);
            return NOA_HIP_ERROR_UNSUPPORTED;
        };
        if (in_dtype == 3 && out_dtype == 3) return run(Cx<float>{}, Cx<float>{});
        if (in_dtype == 3 && out_dtype == 1) return run(Cx<float>{}, float{});
        if (in_dtype == 1 && out_dtype == 1) return run(float{}, float{});
        if (in_dtype == 4 && out_dtype == 4) return run(Cx<double>{}, Cx<double>{});
        if (in_dtype == 4 && out_dtype == 2) return run(Cx<double>{}, double{});
        if (in_dtype == 2 && out_dtype == 2) return run(double{}, double{});
        if (in_dtype == 0 && out_dtype == 1) return run(float{}, float{});
        if (in_dtype == 0 && out_dtype == 2) return run(double{}, double{});
        if (in_dtype == 0 && out_dtype == 3) return run(float{}, Cx<float>{});
        if (in_dtype == 0 && out_dtype == 4) return run(double{}, Cx<double>{});
        g_error = "oracle: unsupported CTF value types";
        return NOA_HIP_ERROR_UNSUPPORTED;
    };
    if (isotropic) return ctf_f64 ? go(double{}, CtfIso<double>{}) : go(float{}, CtfIso<float>{});
    if (ndim4(sh) != 2) { g_error = "oracle: anisotropic CTFs are 2D"; return NOA_HIP_ERROR_UNSUPPORTED; }
    return ctf_f64 ? go(double{}, CtfAniso<double>{}) : go(float{}, CtfAniso<float>{});
}

/* interpolation_weights of a window of 4 to 8 taps (CUBIC, CUBIC_BSPLINE, LANCZOS4/6/8) at fraction f; returns the
 * window size. */
int oracle_window_weights_f64(int interp, double f, double* out) {
    double w[8];
    oracle::window_weights<double>(interp, f, w);
    const int n = oracle::window_size(interp);
    for (int i = 0; i < n; ++i) out[i] = w[i];
    return n;
}
int oracle_window_weights_f32(int interp, float f, float* out) {
    float w[8];
    oracle::window_weights<float>(interp, f, w);
    const int n = oracle::window_size(interp);
    for (int i = 0; i < n; ++i) out[i] = w[i];
    return n;
}

/* Pinned helpers, exposed for the oracle's own tests. */
long long oracle_index2frequency(int centered, long long index, long long size) {
    return oracle::index2frequency(centered != 0, index, size);
}
long long oracle_frequency2index(int centered, long long frequency, long long size) {
    return oracle::frequency2index(centered != 0, frequency, size);
}
double oracle_windowed_sinc_f64(double f, double fs, double fb) { return oracle::windowed_sinc<double>(f, fs, fb); }
float oracle_windowed_sinc_f32(float f, float fs, float fb) { return oracle::windowed_sinc<float>(f, fs, fb); }
long long oracle_blackman_window_size(double fb, double n) { return oracle::blackman_window_size<double>(fb, n); }
int oracle_iwise_threads(long long n, int n_threads) { return oracle::iwise_threads(n, n_threads); }

/* interpolate_spectrum on host arrays (c32 2D/3D, LINEAR or NEAREST), for the interpolation pins. */
void oracle_interpolate_spectrum_3d_c32(const float* data, const long long strides[4], long long nz, long long ny,
                                        long long nx, int centered, int interp, const float* freqs, long long n,
                                        float* out) {
    oracle::Reader3<oracle::Cx<float>> r;
    r.p = reinterpret_cast<const oracle::Cx<float>*>(data);
    r.s0 = 0; r.s1 = strides[1]; r.s2 = strides[2]; r.s3 = strides[3];
    for (long long i = 0; i < n; ++i) {
        auto v = oracle::interpolate_spectrum_3d<oracle::Cx<float>, float>(
            r, freqs[3 * i], freqs[3 * i + 1], freqs[3 * i + 2], nz, ny, nx, centered != 0, interp);
        out[2 * i] = v.re; out[2 * i + 1] = v.im;
    }
}
void oracle_interpolate_spectrum_2d_c32(const float* data, const long long strides[3], long long batch,
                                        long long ny, long long nx, int centered, int interp, const float* freqs,
                                        long long n, float* out) {
    oracle::Reader3<oracle::Cx<float>> r;
    r.p = reinterpret_cast<const oracle::Cx<float>*>(data);
    r.s0 = strides[0]; r.s1 = strides[1]; r.s2 = strides[2];
    for (long long i = 0; i < n; ++i) {
        auto v = oracle::interpolate_spectrum_2d<oracle::Cx<float>, float>(
            r, batch, freqs[2 * i], freqs[2 * i + 1], ny, nx, centered != 0, interp);
        out[2 * i] = v.re; out[2 * i + 1] = v.im;
    }
}

} // extern "C"
