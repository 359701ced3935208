// Decoding and encoding of MRC data on the GPU: noa::io::decode / encode (src/noa/io/Encoding.cpp:33-160) for the
// data types an MRC file stores, between a chunk of raw file bytes in device memory and a strided f32/f64/c32/c64
// device array. One lane per element (per byte for 4-bit data), elements in the file's BDHW order.
//
// Rules (noa's): decoding swaps the bytes of big-endian files, then converts (clamp_cast when clamping; for the
// floating-point outputs used here every stored value is representable, so both are plain conversions; real
// values into a complex output give {value, 0}). Encoding converts with clamp_cast (base/ClampCast.hpp): a float to
// an integer type is 0 for NaN, saturates at the type's limits and otherwise truncates; f64 to f32 or f16 clamps
// to the finite range first; 4-bit values are rounded (half away from zero), clamped to [0, 15] and packed low
// nibble first (Encoding.cpp:138-160).
#include <hip/hip_fp16.h>

#include "common.hpp"
#include "kernels.hpp"

namespace noa_hip {

namespace {

using u16 = uint16_t;

__device__ __forceinline__ u16 bswap16(u16 v) { return static_cast<u16>((v >> 8) | (v << 8)); }
__device__ __forceinline__ u32 bswap32(u32 v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// The stored value at element i of the chunk, as (real, imag) in R.
template<class R>
__device__ __forceinline__ void load_raw(const unsigned char* raw, int dtype, bool swap, i64 i, R& re, R& im) {
    im = R(0);
    switch (dtype) {
        case NOA_HIP_IO_I8: re = static_cast<R>(static_cast<signed char>(raw[i])); break;
        case NOA_HIP_IO_U8: re = static_cast<R>(raw[i]); break;
        case NOA_HIP_IO_I16: case NOA_HIP_IO_U16: case NOA_HIP_IO_F16: {
            u16 v = static_cast<u16>(raw[2 * i] | (raw[2 * i + 1] << 8));
            if (swap) v = bswap16(v);
            if (dtype == NOA_HIP_IO_I16) re = static_cast<R>(static_cast<short>(v));
            else if (dtype == NOA_HIP_IO_U16) re = static_cast<R>(v);
            else re = static_cast<R>(__half2float(__ushort_as_half(v)));
            break;
        }
        case NOA_HIP_IO_F32: {
            u32 v = *reinterpret_cast<const u32*>(raw + 4 * i);
            if (swap) v = bswap32(v);
            re = static_cast<R>(__uint_as_float(v));
            break;
        }
        case NOA_HIP_IO_CI16: {
            u16 a = static_cast<u16>(raw[4 * i] | (raw[4 * i + 1] << 8));
            u16 b = static_cast<u16>(raw[4 * i + 2] | (raw[4 * i + 3] << 8));
            if (swap) { a = bswap16(a); b = bswap16(b); }
            re = static_cast<R>(static_cast<short>(a));
            im = static_cast<R>(static_cast<short>(b));
            break;
        }
        case NOA_HIP_IO_C32: {
            u32 a = *reinterpret_cast<const u32*>(raw + 8 * i);
            u32 b = *reinterpret_cast<const u32*>(raw + 8 * i + 4);
            if (swap) { a = bswap32(a); b = bswap32(b); }
            re = static_cast<R>(__uint_as_float(a));
            im = static_cast<R>(__uint_as_float(b));
            break;
        }
        case NOA_HIP_IO_U4: {
            const unsigned char byte = raw[i / 2];
            re = static_cast<R>((i & 1) ? ((byte >> 4) & 0x0f) : (byte & 0x0f));
            break;
        }
        default: re = R(0); break;
    }
}

template<class Out>
__global__ __launch_bounds__(256) void mrc_decode_kernel(const unsigned char* __restrict__ raw, int dtype, bool swap,
                                                         i64 first, i64 count, Out* __restrict__ out, i64 s0, i64 s1,
                                                         i64 s2, i64 s3, i64 n1, i64 n2, i64 n3) {
    using R = typename vt<Out>::real;
    const i64 i = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= count) return;
    R re, im;
    load_raw<R>(raw, dtype, swap, i, re, im);
    const i64 e = first + i;
    const i64 x = e % n3, r = e / n3;
    const i64 y = r % n2, r2 = r / n2;
    const i64 z = r2 % n1, b = r2 / n1;
    Out* o = out + (b * s0 + z * s1 + y * s2 + x * s3);
    if constexpr (vt<Out>::cx) { Out v; v.x = re; v.y = im; *o = v; }
    else *o = re;
}

// clamp_cast from a float or double to an integer type of limits [lo, hi] (base/ClampCast.hpp): NaN -> 0,
// saturation, truncation. Without clamping: the plain conversion (through a 32-bit integer).
template<class R>
__device__ __forceinline__ int to_int(R v, bool clamp, int lo, int hi) {
    if (clamp) {
        if (v != v) return 0;
        if (v <= static_cast<R>(lo)) return lo;
        if (v >= static_cast<R>(hi)) return hi;
        return static_cast<int>(v);
    }
    return static_cast<int>(v);
}

__device__ __forceinline__ u16 to_half_bits(float v, bool clamp) {
    if (clamp && v == v) v = fminf(fmaxf(v, -65504.f), 65504.f);
    return __half_as_ushort(__float2half_rn(v));
}
__device__ __forceinline__ u16 to_half_bits(double v, bool clamp) {
    if (clamp && v == v) v = fmin(fmax(v, -65504.0), 65504.0);
    const _Float16 h = static_cast<_Float16>(v);   // one rounding from f64
    u16 bits;
    __builtin_memcpy(&bits, &h, 2);
    return bits;
}

template<class R>
__device__ __forceinline__ float to_f32(R v, bool clamp) {
    if constexpr (sizeof(R) == 8) {
        if (clamp && v == v) v = fmin(fmax(v, -3.4028234663852886e38), 3.4028234663852886e38);
    }
    return static_cast<float>(v);
}

// clamp_cast<u32> (NaN -> 0, negative -> 0, saturation), then the 4-bit clamp (Encoding.cpp:152-155).
template<class R>
__device__ __forceinline__ u32 to_u4(R v) {
    const R rv = round(v);
    u32 u;
    if (rv != rv || rv <= R(0)) u = 0;
    else if (rv >= static_cast<R>(4294967295.0)) u = 4294967295u;
    else u = static_cast<u32>(rv);
    return u > 15u ? 15u : u;
}

template<class In>
__device__ __forceinline__ void read_in(const In* p, typename vt<In>::real& re, typename vt<In>::real& im) {
    if constexpr (vt<In>::cx) { const In v = *p; re = v.x; im = v.y; }
    else { re = *p; im = 0; }
}

template<class In>
__global__ __launch_bounds__(256) void mrc_encode_kernel(const In* __restrict__ in, i64 s0, i64 s1, i64 s2, i64 s3,
                                                         i64 n1, i64 n2, i64 n3, i64 first, i64 count, int dtype,
                                                         bool clamp, unsigned char* __restrict__ raw) {
    using R = typename vt<In>::real;
    const i64 i = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    auto at = [&](i64 e) {
        const i64 x = e % n3, r = e / n3;
        const i64 y = r % n2, r2 = r / n2;
        const i64 z = r2 % n1, b = r2 / n1;
        return in + (b * s0 + z * s1 + y * s2 + x * s3);
    };
    if (dtype == NOA_HIP_IO_U4) {
        // One lane per byte: elements 2i and 2i+1 of the chunk, low nibble first.
        if (2 * i >= count) return;
        R a, b, ia, ib;
        read_in(at(first + 2 * i), a, ia);
        read_in(at(first + 2 * i + 1), b, ib);
        raw[i] = static_cast<unsigned char>(to_u4(a) | (to_u4(b) << 4));
        return;
    }
    if (i >= count) return;
    R re, im;
    read_in(at(first + i), re, im);
    switch (dtype) {
        case NOA_HIP_IO_I8: raw[i] = static_cast<unsigned char>(static_cast<signed char>(to_int(re, clamp, -128, 127))); break;
        case NOA_HIP_IO_U8: raw[i] = static_cast<unsigned char>(to_int(re, clamp, 0, 255)); break;
        case NOA_HIP_IO_I16: case NOA_HIP_IO_U16: case NOA_HIP_IO_F16: {
            u16 v;
            if (dtype == NOA_HIP_IO_I16) v = static_cast<u16>(static_cast<short>(to_int(re, clamp, -32768, 32767)));
            else if (dtype == NOA_HIP_IO_U16) v = static_cast<u16>(to_int(re, clamp, 0, 65535));
            else v = to_half_bits(re, clamp);
            raw[2 * i] = static_cast<unsigned char>(v & 0xff);
            raw[2 * i + 1] = static_cast<unsigned char>(v >> 8);
            break;
        }
        case NOA_HIP_IO_F32: *reinterpret_cast<float*>(raw + 4 * i) = to_f32(re, clamp); break;
        case NOA_HIP_IO_CI16: {
            const u16 a = static_cast<u16>(static_cast<short>(to_int(re, clamp, -32768, 32767)));
            const u16 b = static_cast<u16>(static_cast<short>(to_int(im, clamp, -32768, 32767)));
            raw[4 * i] = static_cast<unsigned char>(a & 0xff);
            raw[4 * i + 1] = static_cast<unsigned char>(a >> 8);
            raw[4 * i + 2] = static_cast<unsigned char>(b & 0xff);
            raw[4 * i + 3] = static_cast<unsigned char>(b >> 8);
            break;
        }
        case NOA_HIP_IO_C32:
            *reinterpret_cast<float*>(raw + 8 * i) = to_f32(re, clamp);
            *reinterpret_cast<float*>(raw + 8 * i + 4) = to_f32(im, clamp);
            break;
        default: break;
    }
}

int launch_error(std::string& err) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

} // namespace

int launch_mrc_decode(const MrcChunk& c, hipStream_t stream, std::string& err) {
    if (c.count <= 0) return NOA_HIP_OK;
    const dim3 grid(static_cast<u32>((c.count + 255) / 256)), block(256);
    const auto* raw = static_cast<const unsigned char*>(c.raw);
    switch (c.array_dtype) {
        case NOA_HIP_F32: hipLaunchKernelGGL((mrc_decode_kernel<float>), grid, block, 0, stream, raw, c.file_dtype, c.swap, c.first, c.count, static_cast<float*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3]); break;
        case NOA_HIP_F64: hipLaunchKernelGGL((mrc_decode_kernel<double>), grid, block, 0, stream, raw, c.file_dtype, c.swap, c.first, c.count, static_cast<double*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3]); break;
        case NOA_HIP_C32: hipLaunchKernelGGL((mrc_decode_kernel<float2>), grid, block, 0, stream, raw, c.file_dtype, c.swap, c.first, c.count, static_cast<float2*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3]); break;
        case NOA_HIP_C64: hipLaunchKernelGGL((mrc_decode_kernel<double2>), grid, block, 0, stream, raw, c.file_dtype, c.swap, c.first, c.count, static_cast<double2*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3]); break;
        default: err = "mrc: unsupported array type"; return NOA_HIP_ERROR_UNSUPPORTED;
    }
    return launch_error(err);
}

int launch_mrc_encode(const MrcChunk& c, hipStream_t stream, std::string& err) {
    if (c.count <= 0) return NOA_HIP_OK;
    const i64 lanes = c.file_dtype == NOA_HIP_IO_U4 ? (c.count + 1) / 2 : c.count;
    const dim3 grid(static_cast<u32>((lanes + 255) / 256)), block(256);
    auto* raw = static_cast<unsigned char*>(c.raw);
    switch (c.array_dtype) {
        case NOA_HIP_F32: hipLaunchKernelGGL((mrc_encode_kernel<float>), grid, block, 0, stream, static_cast<const float*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3], c.first, c.count, c.file_dtype, c.clamp, raw); break;
        case NOA_HIP_F64: hipLaunchKernelGGL((mrc_encode_kernel<double>), grid, block, 0, stream, static_cast<const double*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3], c.first, c.count, c.file_dtype, c.clamp, raw); break;
        case NOA_HIP_C32: hipLaunchKernelGGL((mrc_encode_kernel<float2>), grid, block, 0, stream, static_cast<const float2*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3], c.first, c.count, c.file_dtype, c.clamp, raw); break;
        case NOA_HIP_C64: hipLaunchKernelGGL((mrc_encode_kernel<double2>), grid, block, 0, stream, static_cast<const double2*>(c.array), c.s[0], c.s[1], c.s[2], c.s[3], c.n[1], c.n[2], c.n[3], c.first, c.count, c.file_dtype, c.clamp, raw); break;
        default: err = "mrc: unsupported array type"; return NOA_HIP_ERROR_UNSUPPORTED;
    }
    return launch_error(err);
}

} // namespace noa_hip
