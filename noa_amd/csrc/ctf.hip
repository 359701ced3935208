// noa::signal::ctf_anisotropic for 2D rfft layouts (src/noa/signal/CTF.hpp:14-125 operator, :430-500 public
// function) and the CTFAnisotropic<T> model it evaluates (src/noa/signal/core/CTF.hpp:291-575), plus a fused
// form that writes input * CTF and CTF^2 in one pass: the prologue of the CTF-weighted insertion, where noa calls
// the operator twice (once on the slices, once to write the squared CTF as the weights). Then the general
// operator: ctf_isotropic over 1, 2 or 3 dimensions and both CTFs on full layouts.
//
// Elementwise and HBM-bound: one lane per element. The CTF is evaluated once per element, in the CTF's precision,
// with noa's expressions in noa's order.
#include <cstring>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace noa_hip {

// The members of CTFAnisotropic<T> in declaration order (core/CTF.hpp:564-575): the layout of noa's objects, so a
// device array of noa CTFs is read as is.
template<class T>
struct alignas(2 * sizeof(T)) CtfAniso {
    T pixel_size[2];
    T defocus_value, defocus_astigmatism, defocus_angle;   // angstroms (underfocus negative), radians
    T phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
};
static_assert(sizeof(CtfAniso<float>) == 56 && sizeof(CtfAniso<double>) == 112, "noa's CTFAnisotropic layout");

// CTFAnisotropic::value_at (core/CTF.hpp:473-488) with phi2defocus_ (:540-543), cartesian2phi<false> and
// cartesian2rho (xform/core/Polar.hpp:100-135).
template<class T>
__device__ __forceinline__ T ctf_value_at(const CtfAniso<T>& c, T fy, T fx) {
    const T sy = fy / c.pixel_size[0], sx = fx / c.pixel_size[1];
    const T phi = atan2(sy, sx);
    const T rho = sqrt(sy * sy + sx * sx);
    const T r2 = rho * rho;
    const T r4 = r2 * r2;
    const T defocus = c.defocus_value + c.defocus_astigmatism * cos(2 * (phi - c.defocus_angle));
    const T phase = c.k1 * r2 * defocus + c.k2 * r4 - c.phase_shift - c.k3;
    T ctf = -sin(phase);
    if (c.bfactor_forth != 0) ctf *= exp(c.bfactor_forth * r2);
    return ctf * c.scale;
}

template<class T, class In, class Out, class W>
struct CtfParams {
    const In* in; i64 is_b; i32 is_y, is_x;     // null: the CTF itself is written
    Out* out; i64 os_b; i32 os_y, os_x;
    W* wout; i64 ws_b; i32 ws_y, ws_x;          // fused form: CTF^2 (null otherwise)
    const CtfAniso<T>* ctfs;                    // batched (device), or null: `single`
    CtfAniso<T> single;
    u32 h, wr, n;                               // logical height, rfft width, elements (batch * h * wr)
    T step[2];                                  // fftfreq per frequency along y and x
    bool src_centered, dst_centered, ctf_abs, ctf_squared;
};

// Elements per lane: each block covers CTF_EPT * 256 consecutive elements (fewer, longer-lived blocks than one
// element per lane, whose dispatch rate bounded the launch).
constexpr u32 CTF_EPT = 4;

// Element t (x fastest: loads and stores are contiguous along the rows).
template<class T, class In, class Out, class W, bool HAS_INPUT>
__device__ __forceinline__ void ctf_element(const CtfParams<T, In, Out, W>& p, u32 t) {
    using RIn = typename vt<In>::real;
    const u32 row = t / p.wr, x = t - row * p.wr;
    const u32 b = row / p.h, y = row - b * p.h;
    const CtfAniso<T> c = p.ctfs ? p.ctfs[b] : p.single;
    const i32 fy = index2frequency(p.dst_centered, static_cast<i32>(y), static_cast<i32>(p.h));
    // fftfreq = frequency * step (signal/CTF.hpp:84-85), then value_at in the CTF's precision.
    T ctf = ctf_value_at(c, static_cast<T>(fy) * p.step[0], static_cast<T>(x) * p.step[1]);
    if constexpr (HAS_INPUT) {
        if (p.wout) {
            // Fused form: the value is multiplied by the CTF, the weight is the squared CTF.
            const i32 iy = frequency2index(p.src_centered, fy, static_cast<i32>(p.h));   // remap_indices<REMAP, true>
            const In v = p.in[static_cast<i64>(b) * p.is_b + static_cast<i64>(iy) * p.is_y + static_cast<i64>(x) * p.is_x];
            p.out[static_cast<i64>(b) * p.os_b + static_cast<i64>(y) * p.os_y + static_cast<i64>(x) * p.os_x] =
                vcast<Out>(vmul(v, static_cast<RIn>(ctf)));
            p.wout[static_cast<i64>(b) * p.ws_b + static_cast<i64>(y) * p.ws_y + static_cast<i64>(x) * p.ws_x] =
                static_cast<W>(ctf * ctf);
            return;
        }
    }
    if (p.ctf_squared) ctf *= ctf;
    else if (p.ctf_abs) ctf = fabs(ctf);
    Out* const o = p.out + (static_cast<i64>(b) * p.os_b + static_cast<i64>(y) * p.os_y + static_cast<i64>(x) * p.os_x);
    if constexpr (HAS_INPUT) {
        const i32 iy = frequency2index(p.src_centered, fy, static_cast<i32>(p.h));
        const In v = p.in[static_cast<i64>(b) * p.is_b + static_cast<i64>(iy) * p.is_y + static_cast<i64>(x) * p.is_x];
        *o = vcast<Out>(vmul(v, static_cast<RIn>(ctf)));
    } else {
        Out value = vzero<Out>();
        if constexpr (vt<Out>::cx) value.x = static_cast<typename vt<Out>::real>(ctf);
        else value = static_cast<Out>(ctf);
        *o = value;
    }
}

template<class T, class In, class Out, class W, bool HAS_INPUT>
__global__ __launch_bounds__(256) void ctf_kernel(const CtfParams<T, In, Out, W> p) {
    const u32 base = blockIdx.x * (256u * CTF_EPT) + threadIdx.x;
#pragma unroll
    for (u32 e = 0; e < CTF_EPT; ++e) {
        const u32 t = base + e * 256u;
        if (t < p.n) ctf_element<T, In, Out, W, HAS_INPUT>(p, t);
    }
}

// The fused form (value x CTF and CTF^2) with the element loads of a lane issued before any CTF is evaluated, so
// that they are in flight while the CTFs are computed; the same expressions as ctf_element.
constexpr u32 CTF_FUSED_EPT = CTF_EPT;   // 8 measured slower (fewer resident waves)
template<class T, class In, class Out, class W>
__global__ __launch_bounds__(256) void ctf_fused_kernel(const CtfParams<T, In, Out, W> p) {
    using RIn = typename vt<In>::real;
    const u32 base = blockIdx.x * (256u * CTF_FUSED_EPT) + threadIdx.x;
    In v[CTF_FUSED_EPT];
    u32 bs[CTF_FUSED_EPT], ys[CTF_FUSED_EPT], xs[CTF_FUSED_EPT];
    i32 fys[CTF_FUSED_EPT];
#pragma unroll
    for (u32 e = 0; e < CTF_FUSED_EPT; ++e) {
        const u32 t = min(base + e * 256u, p.n - 1);      // lanes past the end repeat the last element, not stored
        const u32 row = t / p.wr, x = t - row * p.wr;
        const u32 b = row / p.h, y = row - b * p.h;
        const i32 fy = index2frequency(p.dst_centered, static_cast<i32>(y), static_cast<i32>(p.h));
        const i32 iy = frequency2index(p.src_centered, fy, static_cast<i32>(p.h));   // remap_indices<REMAP, true>
        v[e] = p.in[static_cast<i64>(b) * p.is_b + static_cast<i64>(iy) * p.is_y + static_cast<i64>(x) * p.is_x];
        bs[e] = b; ys[e] = y; xs[e] = x; fys[e] = fy;
    }
#pragma unroll
    for (u32 e = 0; e < CTF_FUSED_EPT; ++e) {
        if (base + e * 256u >= p.n) continue;
        const u32 b = bs[e], y = ys[e], x = xs[e];
        const CtfAniso<T> c = p.ctfs ? p.ctfs[b] : p.single;
        // fftfreq = frequency * step (signal/CTF.hpp:84-85), then value_at in the CTF's precision.
        const T ctf = ctf_value_at(c, static_cast<T>(fys[e]) * p.step[0], static_cast<T>(x) * p.step[1]);
        p.out[static_cast<i64>(b) * p.os_b + static_cast<i64>(y) * p.os_y + static_cast<i64>(x) * p.os_x] =
            vcast<Out>(vmul(v[e], static_cast<RIn>(ctf)));
        p.wout[static_cast<i64>(b) * p.ws_b + static_cast<i64>(y) * p.ws_y + static_cast<i64>(x) * p.ws_x] =
            static_cast<W>(ctf * ctf);
    }
}

namespace {

// Linspace{start, stop <= 0 ? highest_fftfreq(n) : stop, endpoint}.for_size(n / 2 + 1).step (signal/CTF.hpp:65-77,
// runtime/core/Iwise.hpp:67-78, fft/core/Frequency.hpp:9-20), in the coordinate precision.
template<class T>
T fftfreq_step(i64 n, double start, double stop, bool endpoint) {
    const i64 size = n / 2 + 1;
    const T t_start = static_cast<T>(start), t_stop_opt = static_cast<T>(stop);
    const T t_stop = t_stop_opt <= 0 ? static_cast<T>(n / 2) / static_cast<T>(n) : t_stop_opt;
    const i64 count = std::max<i64>(1, size - (endpoint ? 1 : 0));
    return (t_stop - t_start) / static_cast<T>(count);
}

template<class T, class In, class Out, class W>
int run(const CtfDesc& d, hipStream_t stream, std::string& err) {
    CtfParams<T, In, Out, W> p{};
    p.in = static_cast<const In*>(d.input.data);
    p.is_b = d.input.s0; p.is_y = static_cast<i32>(d.input.s1); p.is_x = static_cast<i32>(d.input.s2);
    p.out = static_cast<Out*>(d.output.data);
    p.os_b = d.output.s0; p.os_y = static_cast<i32>(d.output.s1); p.os_x = static_cast<i32>(d.output.s2);
    p.wout = static_cast<W*>(d.output_weight.data);
    p.ws_b = d.output_weight.s0; p.ws_y = static_cast<i32>(d.output_weight.s1); p.ws_x = static_cast<i32>(d.output_weight.s2);
    p.ctfs = d.ctf_batched ? static_cast<const CtfAniso<T>*>(d.ctf_data) : nullptr;
    if (!d.ctf_batched) std::memcpy(&p.single, d.ctf_single, sizeof(CtfAniso<T>));
    p.h = static_cast<u32>(d.shape[2]);
    p.wr = static_cast<u32>(d.shape[3] / 2 + 1);
    const i64 n = d.shape[0] * d.shape[2] * (d.shape[3] / 2 + 1);
    if (n > 0x7fffffffLL) { err = "ctf: more than 2^31 elements in one call"; return NOA_HIP_ERROR_UNSUPPORTED; }
    p.n = static_cast<u32>(n);
    p.step[0] = fftfreq_step<T>(d.shape[2], d.range_start, d.range_stop, d.endpoint);
    p.step[1] = fftfreq_step<T>(d.shape[3], d.range_start, d.range_stop, d.endpoint);
    p.src_centered = d.src_centered;
    p.dst_centered = d.dst_centered;
    p.ctf_abs = d.ctf_abs;
    p.ctf_squared = d.ctf_squared;
    const i64 blocks = (n + 256 * CTF_EPT - 1) / (256 * CTF_EPT);
    if (blocks == 0) return NOA_HIP_OK;
    const dim3 grid(static_cast<u32>(blocks)), block(256);
    if (p.in && p.wout) {
        const dim3 fgrid(static_cast<u32>((n + 256 * CTF_FUSED_EPT - 1) / (256 * CTF_FUSED_EPT)));
        if constexpr (vt<In>::cx || !vt<Out>::cx) hipLaunchKernelGGL((ctf_fused_kernel<T, In, Out, W>), fgrid, block, 0, stream, p);
    } else if (p.in) {
        if constexpr (vt<In>::cx || !vt<Out>::cx) hipLaunchKernelGGL((ctf_kernel<T, In, Out, W, true>), grid, block, 0, stream, p);
    } else {
        hipLaunchKernelGGL((ctf_kernel<T, In, Out, W, false>), grid, block, 0, stream, p);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class T>
int dispatch(const CtfDesc& d, hipStream_t stream, std::string& err) {
    const int in = d.input.data ? d.input.dtype : NOA_HIP_DTYPE_NONE, out = d.output.dtype;
    switch (in) {
        case NOA_HIP_C32:
            if (out == NOA_HIP_C32) return run<T, float2, float2, float>(d, stream, err);
            if (out == NOA_HIP_F32) return run<T, float2, float, float>(d, stream, err);
            break;
        case NOA_HIP_F32:
            if (out == NOA_HIP_F32) return run<T, float, float, float>(d, stream, err);
            break;
        case NOA_HIP_C64:
            if (out == NOA_HIP_C64) return run<T, double2, double2, double>(d, stream, err);
            if (out == NOA_HIP_F64) return run<T, double2, double, double>(d, stream, err);
            break;
        case NOA_HIP_F64:
            if (out == NOA_HIP_F64) return run<T, double, double, double>(d, stream, err);
            break;
        default:   // no input: the CTF is written as a real or complex value
            if (out == NOA_HIP_F32) return run<T, float, float, float>(d, stream, err);
            if (out == NOA_HIP_C32) return run<T, float, float2, float>(d, stream, err);
            if (out == NOA_HIP_F64) return run<T, double, double, double>(d, stream, err);
            if (out == NOA_HIP_C64) return run<T, double, double2, double>(d, stream, err);
            break;
    }
    err = "ctf: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_ctf(const CtfDesc& d, hipStream_t stream, std::string& err) {
    return d.ctf_f64 ? dispatch<double>(d, stream, err) : dispatch<float>(d, stream, err);
}

// ------------------------------------------------------------------------------------------------------------------
// The general operator (signal/CTF.hpp:14-125) over 1, 2 or 3 dimensions, half or full layouts, isotropic
// (CTFIsotropic, core/CTF.hpp:30-272) or anisotropic CTFs: ctf_isotropic (:216-421) and the full layouts of
// ctf_anisotropic (:430-500). The host has already put the dimensions in noa's order.
// ------------------------------------------------------------------------------------------------------------------

// The members of CTFIsotropic<T> in declaration order (core/CTF.hpp:259-271).
template<class T>
struct CtfIso {
    T pixel_size, defocus, phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
};
static_assert(sizeof(CtfIso<float>) == 44 && sizeof(CtfIso<double>) == 88, "noa's CTFIsotropic layout");

// CTFIsotropic::value_at (core/CTF.hpp:204-212).
template<class T>
__device__ __forceinline__ T ctf_value_at(const CtfIso<T>& c, T fftfreq) {
    const T r1 = fftfreq / c.pixel_size;
    const T r2 = r1 * r1;
    const T r4 = r2 * r2;
    const T phase = (c.k1 * r2 * c.defocus) + (c.k2 * r4) - c.phase_shift - c.k3;
    T ctf = -sin(phase);
    if (c.bfactor_forth != 0) ctf *= exp(c.bfactor_forth * r2);
    return ctf * c.scale;
}

template<class T, class C, class In, class Out>
struct CtfNdParams {
    const In* in; i64 is[4];
    Out* out; i64 os[4];
    const C* ctfs;                  // batched (device), or null: `single`
    C single;
    i32 dims[3];                    // logical sizes of the iterated dimensions, outermost first
    u32 iter[3];                    // iterated extents (the last one halved for a half layout)
    u32 chunks;                     // blocks per innermost line
    T step[3];
    T start;
    bool half, src_centered, dst_centered, ctf_abs, ctf_squared;
};

// One block of 256 lanes per stretch of CTF_EPT * 256 elements of an innermost line: the line's outer indices are
// decoded once per block, the lanes run along the innermost dimension (contiguous for a row-major output).
template<int N, bool HAS_INPUT, class T, class C, class In, class Out>
__global__ __launch_bounds__(256) void ctf_nd_kernel(const CtfNdParams<T, C, In, Out> p) {
    constexpr bool ISO = !std::is_same<C, CtfAniso<T>>::value;
    using RIn = typename vt<In>::real;
    const u32 line = blockIdx.x / p.chunks;
    const u32 x0 = (blockIdx.x - line * p.chunks) * (256u * CTF_EPT) + threadIdx.x;
    u32 idx[3];
    u32 r = line;
    for (int i = N - 2; i >= 0; --i) {
        idx[i] = r % p.iter[i];
        r /= p.iter[i];
    }
    const u32 b = r;
    const C c = p.ctfs ? p.ctfs[b] : p.single;
#pragma unroll
    for (u32 e = 0; e < CTF_EPT; ++e) {
        const u32 x = x0 + e * 256u;
        if (x >= p.iter[N - 1]) return;
        idx[N - 1] = x;
        T f[3];
        i64 in_off = static_cast<i64>(b) * p.is[0], out_off = static_cast<i64>(b) * p.os[0];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const bool halfdim = p.half && i == N - 1;
            const i32 j = static_cast<i32>(idx[i]);
            const i32 freq = halfdim ? j : index2frequency(p.dst_centered, j, p.dims[i]);
            const i32 src = halfdim ? j : frequency2index(p.src_centered, freq, p.dims[i]);   // remap_indices
            f[i] = static_cast<T>(freq) * p.step[i];
            if constexpr (HAS_INPUT) in_off += static_cast<i64>(src) * p.is[i + 1];
            out_off += static_cast<i64>(j) * p.os[i + 1];
        }
        T ctf;
        if constexpr (ISO) {
            T arg;
            if constexpr (N == 1) arg = f[0] + p.start;
            else if constexpr (N == 2) arg = sqrt(f[0] * f[0] + f[1] * f[1]);
            else arg = sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
            ctf = ctf_value_at(c, arg);
        } else {
            ctf = ctf_value_at(c, f[0], f[N - 1]);
        }
        if (p.ctf_squared) ctf *= ctf;
        else if (p.ctf_abs) ctf = fabs(ctf);
        if constexpr (HAS_INPUT) {
            p.out[out_off] = vcast<Out>(vmul(p.in[in_off], static_cast<RIn>(ctf)));
        } else {
            Out value = vzero<Out>();
            if constexpr (vt<Out>::cx) value.x = static_cast<typename vt<Out>::real>(ctf);
            else value = static_cast<Out>(ctf);
            p.out[out_off] = value;
        }
    }
}

// The same over the flattened elements (lines shorter than 256 * CTF_EPT leave most of the per-line kernel's lanes
// idle): CTF_EPT elements per lane, the input loads issued before any CTF is evaluated; for fewer than 2^32 elements.
template<int N, bool HAS_INPUT, class T, class C, class In, class Out>
__global__ __launch_bounds__(256) void ctf_nd_flat_kernel(const CtfNdParams<T, C, In, Out> p, u32 total) {
    constexpr bool ISO = !std::is_same<C, CtfAniso<T>>::value;
    using RIn = typename vt<In>::real;
    const u32 base = blockIdx.x * (256u * CTF_EPT) + threadIdx.x;
    In v[CTF_EPT];
    i64 oo[CTF_EPT];
    T fs[CTF_EPT][3];
    u32 bs[CTF_EPT];
#pragma unroll
    for (u32 e = 0; e < CTF_EPT; ++e) {
        const u32 t = min(base + e * 256u, total - 1);   // lanes past the end repeat the last element, not stored
        u32 idx[3];
        const u32 line = t / p.iter[N - 1];
        idx[N - 1] = t - line * p.iter[N - 1];
        u32 r = line;
#pragma unroll
        for (int i = N - 2; i >= 0; --i) {
            idx[i] = r % p.iter[i];
            r /= p.iter[i];
        }
        const u32 b = r;
        i64 in_off = static_cast<i64>(b) * p.is[0], out_off = static_cast<i64>(b) * p.os[0];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const bool halfdim = p.half && i == N - 1;
            const i32 j = static_cast<i32>(idx[i]);
            const i32 freq = halfdim ? j : index2frequency(p.dst_centered, j, p.dims[i]);
            const i32 src = halfdim ? j : frequency2index(p.src_centered, freq, p.dims[i]);   // remap_indices
            fs[e][i] = static_cast<T>(freq) * p.step[i];
            if constexpr (HAS_INPUT) in_off += static_cast<i64>(src) * p.is[i + 1];
            out_off += static_cast<i64>(j) * p.os[i + 1];
        }
        if constexpr (HAS_INPUT) v[e] = p.in[in_off];
        oo[e] = out_off;
        bs[e] = b;
    }
#pragma unroll
    for (u32 e = 0; e < CTF_EPT; ++e) {
        if (base + e * 256u >= total) continue;
        const C c = p.ctfs ? p.ctfs[bs[e]] : p.single;
        const T* f = fs[e];
        T ctf;
        if constexpr (ISO) {
            T arg;
            if constexpr (N == 1) arg = f[0] + p.start;
            else if constexpr (N == 2) arg = sqrt(f[0] * f[0] + f[1] * f[1]);
            else arg = sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
            ctf = ctf_value_at(c, arg);
        } else {
            ctf = ctf_value_at(c, f[0], f[N - 1]);
        }
        if (p.ctf_squared) ctf *= ctf;
        else if (p.ctf_abs) ctf = fabs(ctf);
        if constexpr (HAS_INPUT) {
            p.out[oo[e]] = vcast<Out>(vmul(v[e], static_cast<RIn>(ctf)));
        } else {
            Out value = vzero<Out>();
            if constexpr (vt<Out>::cx) value.x = static_cast<typename vt<Out>::real>(ctf);
            else value = static_cast<Out>(ctf);
            p.out[oo[e]] = value;
        }
    }
}

namespace {

template<class T, class C, class In, class Out>
int run_nd(const CtfNdDesc& d, hipStream_t stream, std::string& err) {
    CtfNdParams<T, C, In, Out> p{};
    p.in = static_cast<const In*>(d.input.data);
    p.out = static_cast<Out*>(d.output.data);
    for (int i = 0; i < 4; ++i) { p.is[i] = d.input.s[i]; p.os[i] = d.output.s[i]; }
    p.ctfs = d.ctf_batched ? static_cast<const C*>(d.ctf_data) : nullptr;
    if (!d.ctf_batched) std::memcpy(&p.single, d.ctf_single, sizeof(C));
    const int n = d.n_dims;
    i64 lines = d.batch;
    for (int i = 0; i < n; ++i) {
        const i64 it = (d.half && i == n - 1) ? d.dims[i] / 2 + 1 : d.dims[i];
        if (d.dims[i] > 0x7fffffffLL) { err = "ctf: a dimension exceeds 2^31"; return NOA_HIP_ERROR_UNSUPPORTED; }
        p.dims[i] = static_cast<i32>(d.dims[i]);
        p.iter[i] = static_cast<u32>(it);
        p.step[i] = fftfreq_step<T>(d.dims[i], d.range_start, d.range_stop, d.endpoint);
        if (i < n - 1) lines *= it;
    }
    p.start = static_cast<T>(d.range_start);
    p.chunks = (p.iter[n - 1] + 256u * CTF_EPT - 1) / (256u * CTF_EPT);
    const i64 blocks = lines * p.chunks;
    if (blocks == 0) return NOA_HIP_OK;
    if (blocks > 0xffffffffLL) { err = "ctf: more than 2^32 blocks in one call"; return NOA_HIP_ERROR_UNSUPPORTED; }
    p.half = d.half;
    p.src_centered = d.src_centered;
    p.dst_centered = d.dst_centered;
    p.ctf_abs = d.ctf_abs;
    p.ctf_squared = d.ctf_squared;
    const dim3 grid(static_cast<u32>(blocks)), block(256);
    const i64 total = lines * static_cast<i64>(p.iter[n - 1]);
    const bool flat = total < (1LL << 32);
    const dim3 fgrid(static_cast<u32>((total + 256 * CTF_EPT - 1) / (256 * CTF_EPT)));
    auto launch = [&](auto n_tag) {
        constexpr int NN = decltype(n_tag)::value;
        if constexpr (NN == 2 || !std::is_same<C, CtfAniso<T>>::value) {
            if (p.in) {
                if constexpr (vt<In>::cx || !vt<Out>::cx) {
                    if (flat)
                        hipLaunchKernelGGL((ctf_nd_flat_kernel<NN, true, T, C, In, Out>), fgrid, block, 0, stream, p,
                                           static_cast<u32>(total));
                    else
                        hipLaunchKernelGGL((ctf_nd_kernel<NN, true, T, C, In, Out>), grid, block, 0, stream, p);
                }
            } else if (flat) {
                hipLaunchKernelGGL((ctf_nd_flat_kernel<NN, false, T, C, In, Out>), fgrid, block, 0, stream, p,
                                   static_cast<u32>(total));
            } else {
                hipLaunchKernelGGL((ctf_nd_kernel<NN, false, T, C, In, Out>), grid, block, 0, stream, p);
            }
        }
    };
    if (n == 1) launch(std::integral_constant<int, 1>{});
    else if (n == 2) launch(std::integral_constant<int, 2>{});
    else launch(std::integral_constant<int, 3>{});
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class T, class C>
int dispatch_nd(const CtfNdDesc& d, hipStream_t stream, std::string& err) {
    const int in = d.input.data ? d.input.dtype : NOA_HIP_DTYPE_NONE, out = d.output.dtype;
    switch (in) {
        case NOA_HIP_C32:
            if (out == NOA_HIP_C32) return run_nd<T, C, float2, float2>(d, stream, err);
            if (out == NOA_HIP_F32) return run_nd<T, C, float2, float>(d, stream, err);
            break;
        case NOA_HIP_F32:
            if (out == NOA_HIP_F32) return run_nd<T, C, float, float>(d, stream, err);
            break;
        case NOA_HIP_C64:
            if (out == NOA_HIP_C64) return run_nd<T, C, double2, double2>(d, stream, err);
            if (out == NOA_HIP_F64) return run_nd<T, C, double2, double>(d, stream, err);
            break;
        case NOA_HIP_F64:
            if (out == NOA_HIP_F64) return run_nd<T, C, double, double>(d, stream, err);
            break;
        default:
            if (out == NOA_HIP_F32) return run_nd<T, C, float, float>(d, stream, err);
            if (out == NOA_HIP_C32) return run_nd<T, C, float, float2>(d, stream, err);
            if (out == NOA_HIP_F64) return run_nd<T, C, double, double>(d, stream, err);
            if (out == NOA_HIP_C64) return run_nd<T, C, double, double2>(d, stream, err);
            break;
    }
    err = "ctf: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_ctf_nd(const CtfNdDesc& d, hipStream_t stream, std::string& err) {
    if (d.isotropic)
        return d.ctf_f64 ? dispatch_nd<double, CtfIso<double>>(d, stream, err)
                         : dispatch_nd<float, CtfIso<float>>(d, stream, err);
    if (d.n_dims != 2) { err = "ctf_anisotropic: 2D spectra only"; return NOA_HIP_ERROR_UNSUPPORTED; }
    return d.ctf_f64 ? dispatch_nd<double, CtfAniso<double>>(d, stream, err)
                     : dispatch_nd<float, CtfAniso<float>>(d, stream, err);
}

} // namespace noa_hip
