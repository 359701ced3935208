// Elementwise neighbours of the central-slice operators in a reconstruction:
//   - noa::fft::remap between the half-to-half rfft layouts (src/noa/fft/Remap.hpp:205-340; index map
//     fft/core/Frequency.hpp:198-225);
//   - noa::xform::fourier_interpolation_correction (src/noa/xform/FourierProject.hpp:1008-1061, :2062-2087).
// One lane per element, 256-lane workgroups along the innermost dimension (coalesced reads and writes).
#include "common.hpp"
#include "kernels.hpp"

#include <cmath>

namespace noa_hip {

template<class T>
__global__ __launch_bounds__(256) void remap_kernel(const T* in, T* out, i64 is0, i64 is1,
                                                    i64 is2, i64 is3, i64 os0, i64 os1, i64 os2, i64 os3, i32 d,
                                                    i32 h, i32 wr, i32 tiles_x, bool src_c, bool dst_c) {
    const i32 tile = static_cast<i32>(blockIdx.x % static_cast<u32>(tiles_x));
    const i64 row = blockIdx.x / static_cast<u32>(tiles_x);   // (b * d + z) * h + y
    const i32 x = tile * 256 + static_cast<i32>(threadIdx.x);
    if (x >= wr) return;
    const i32 y = static_cast<i32>(row % h);
    const i64 bz = row / h;
    const i32 z = static_cast<i32>(bz % d);
    const i64 b = bz / d;
    const i32 oz = frequency2index(dst_c, index2frequency(src_c, z, d), d);
    const i32 oy = frequency2index(dst_c, index2frequency(src_c, y, h), h);
    out[b * os0 + oz * os1 + oy * os2 + x * os3] = in[b * is0 + z * is1 + y * is2 + x * is3];
}

// The same two operators over the flattened elements, four per lane with every load issued before any store (the
// row-per-block kernels above leave lanes idle on rows shorter than 256 and keep one load in flight per lane); for
// launches below 2^31 elements. Same per-element expressions.
constexpr i32 SPEC_EPT = 4;

template<class T>
__global__ __launch_bounds__(256) void remap_flat_kernel(const T* in, T* out, i64 is0, i64 is1, i64 is2, i64 is3,
                                                         i64 os0, i64 os1, i64 os2, i64 os3, u32 d, u32 h, u32 wr,
                                                         u32 n, bool src_c, bool dst_c) {
    const u32 base = blockIdx.x * (256u * SPEC_EPT) + threadIdx.x;
    T v[SPEC_EPT];
    i64 o[SPEC_EPT];
#pragma unroll
    for (i32 e = 0; e < SPEC_EPT; ++e) {
        const u32 t = min(base + static_cast<u32>(e) * 256u, n - 1);
        const u32 row = t / wr, x = t - row * wr;
        const u32 bz = row / h, y = row - bz * h;
        const u32 b = bz / d, z = bz - b * d;
        const i32 oz = frequency2index(dst_c, index2frequency(src_c, static_cast<i32>(z), static_cast<i32>(d)),
                                       static_cast<i32>(d));
        const i32 oy = frequency2index(dst_c, index2frequency(src_c, static_cast<i32>(y), static_cast<i32>(h)),
                                       static_cast<i32>(h));
        v[e] = in[b * is0 + z * is1 + y * is2 + x * is3];
        o[e] = b * os0 + oz * os1 + oy * os2 + x * os3;
    }
#pragma unroll
    for (i32 e = 0; e < SPEC_EPT; ++e)
        if (base + static_cast<u32>(e) * 256u < n) out[o[e]] = v[e];
}

template<class In, class Out, class C>
__global__ __launch_bounds__(256) void gridding_flat_kernel(const In* in, Out* out, i64 is0, i64 is1, i64 is2,
                                                            i64 is3, i64 os0, i64 os1, i64 os2, i64 os3, u32 d, u32 h,
                                                            u32 w, u32 n, C fs0, C fs1, C fs2, C half0, C half1,
                                                            C half2, bool post) {
    const u32 base = blockIdx.x * (256u * SPEC_EPT) + threadIdx.x;
    In v[SPEC_EPT];
    u32 js[SPEC_EPT], ks[SPEC_EPT], ls[SPEC_EPT], bs[SPEC_EPT];
#pragma unroll
    for (i32 e = 0; e < SPEC_EPT; ++e) {
        const u32 t = min(base + static_cast<u32>(e) * 256u, n - 1);
        const u32 row = t / w, l = t - row * w;
        const u32 bj = row / h, k = row - bj * h;
        const u32 b = bj / d, j = bj - b * d;
        v[e] = in[b * is0 + j * is1 + k * is2 + l * is3];
        js[e] = j; ks[e] = k; ls[e] = l; bs[e] = b;
    }
#pragma unroll
    for (i32 e = 0; e < SPEC_EPT; ++e) {
        if (base + static_cast<u32>(e) * 256u >= n) continue;
        // GriddingCorrection::operator() (:1037-1054).
        C dz = static_cast<C>(static_cast<i32>(js[e])), dy = static_cast<C>(static_cast<i32>(ks[e]));
        C dx = static_cast<C>(static_cast<i32>(ls[e]));
        dz -= half0; dy -= half1; dx -= half2;
        dz /= fs0; dy /= fs1; dx /= fs2;
        const C PI = static_cast<C>(3.141592653589793238462643383279502884);
        const C radius = sqrt((dz * dz + dy * dy) + dx * dx);
        const C xpi = PI * radius;
        const C sinc = xpi == 0 ? C(1) : sin(xpi) / xpi;
        const In sinc2 = static_cast<In>(sinc * sinc);
        out[bs[e] * os0 + js[e] * os1 + ks[e] * os2 + ls[e] * os3] = static_cast<Out>(post ? v[e] / sinc2 : v[e] * sinc2);
    }
}

template<class In, class Out, class C>
__global__ __launch_bounds__(256) void gridding_kernel(const In* in, Out* out, i64 is0, i64 is1, i64 is2,
                                                       i64 is3, i64 os0, i64 os1, i64 os2, i64 os3, i32 d, i32 h,
                                                       i32 w, i32 tiles_x, C fs0, C fs1, C fs2, C half0, C half1,
                                                       C half2, bool post) {
    const i32 tile = static_cast<i32>(blockIdx.x % static_cast<u32>(tiles_x));
    const i64 row = blockIdx.x / static_cast<u32>(tiles_x);
    const i32 l = tile * 256 + static_cast<i32>(threadIdx.x);
    if (l >= w) return;
    const i32 k = static_cast<i32>(row % h);
    const i64 bj = row / h;
    const i32 j = static_cast<i32>(bj % d);
    const i64 b = bj / d;
    // GriddingCorrection::operator() (:1037-1054).
    C dz = static_cast<C>(j), dy = static_cast<C>(k), dx = static_cast<C>(l);
    dz -= half0; dy -= half1; dx -= half2;
    dz /= fs0; dy /= fs1; dx /= fs2;
    const C PI = static_cast<C>(3.141592653589793238462643383279502884);
    const C radius = sqrt((dz * dz + dy * dy) + dx * dx);
    const C xpi = PI * radius;
    const C sinc = xpi == 0 ? C(1) : sin(xpi) / xpi;
    const In sinc2 = static_cast<In>(sinc * sinc);
    const In value = in[b * is0 + j * is1 + k * is2 + l * is3];
    out[b * os0 + j * os1 + k * os2 + l * os3] = static_cast<Out>(post ? value / sinc2 : value * sinc2);
}

namespace {

size_t elem_bytes(int dtype) {
    switch (dtype) {
        case NOA_HIP_F32: case NOA_HIP_C16: case NOA_HIP_CBF16: return 4;
        case NOA_HIP_F64: case NOA_HIP_C32: return 8;
        case NOA_HIP_C64: return 16;
        default: return 0;
    }
}

template<class T>
int run_remap(const RemapDesc& dd, hipStream_t stream, std::string& err) {
    const i32 d = static_cast<i32>(dd.shape[1]), h = static_cast<i32>(dd.shape[2]);
    const i32 wr = static_cast<i32>(dd.shape[3] / 2 + 1);
    const i32 tiles_x = (wr + 255) / 256;
    const i64 blocks = dd.shape[0] * d * static_cast<i64>(h) * tiles_x;
    if (blocks == 0) return NOA_HIP_OK;
    const i64 n = dd.shape[0] * d * static_cast<i64>(h) * wr;
    if (n < (1LL << 31)) {
        hipLaunchKernelGGL((remap_flat_kernel<T>), dim3(static_cast<u32>((n + 256 * SPEC_EPT - 1) / (256 * SPEC_EPT))),
                           dim3(256), 0, stream, static_cast<const T*>(dd.input.data), static_cast<T*>(dd.output.data),
                           dd.input.s[0], dd.input.s[1], dd.input.s[2], dd.input.s[3], dd.output.s[0], dd.output.s[1],
                           dd.output.s[2], dd.output.s[3], static_cast<u32>(d), static_cast<u32>(h),
                           static_cast<u32>(wr), static_cast<u32>(n), dd.src_centered, dd.dst_centered);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
        return NOA_HIP_OK;
    }
    if (blocks > 0xffffffffLL) { err = "remap: too many elements for one launch"; return NOA_HIP_ERROR_UNSUPPORTED; }
    hipLaunchKernelGGL((remap_kernel<T>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                       static_cast<const T*>(dd.input.data), static_cast<T*>(dd.output.data), dd.input.s[0],
                       dd.input.s[1], dd.input.s[2], dd.input.s[3], dd.output.s[0], dd.output.s[1], dd.output.s[2],
                       dd.output.s[3], d, h, wr, tiles_x, dd.src_centered, dd.dst_centered);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class In, class Out, class C>
int run_gridding(const GriddingDesc& g, hipStream_t stream, std::string& err) {
    const i32 d = static_cast<i32>(g.shape[1]), h = static_cast<i32>(g.shape[2]), w = static_cast<i32>(g.shape[3]);
    const i32 tiles_x = (w + 255) / 256;
    const i64 blocks = g.shape[0] * d * static_cast<i64>(h) * tiles_x;
    if (blocks == 0) return NOA_HIP_OK;
    if (blocks > 0xffffffffLL) { err = "fourier_interpolation_correction: too many elements"; return NOA_HIP_ERROR_UNSUPPORTED; }
    // m_f_shape and m_half = shape / 2 * (shape != 1) (:1027-1033).
    const C fs0 = static_cast<C>(d), fs1 = static_cast<C>(h), fs2 = static_cast<C>(w);
    const C h0 = fs0 / 2 * static_cast<C>(d != 1), h1 = fs1 / 2 * static_cast<C>(h != 1), h2 = fs2 / 2 * static_cast<C>(w != 1);
    const i64 n = g.shape[0] * d * static_cast<i64>(h) * w;
    if (n < (1LL << 31)) {
        hipLaunchKernelGGL((gridding_flat_kernel<In, Out, C>),
                           dim3(static_cast<u32>((n + 256 * SPEC_EPT - 1) / (256 * SPEC_EPT))), dim3(256), 0, stream,
                           static_cast<const In*>(g.input.data), static_cast<Out*>(g.output.data), g.input.s[0],
                           g.input.s[1], g.input.s[2], g.input.s[3], g.output.s[0], g.output.s[1], g.output.s[2],
                           g.output.s[3], static_cast<u32>(d), static_cast<u32>(h), static_cast<u32>(w),
                           static_cast<u32>(n), fs0, fs1, fs2, h0, h1, h2, g.post);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
        return NOA_HIP_OK;
    }
    hipLaunchKernelGGL((gridding_kernel<In, Out, C>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                       static_cast<const In*>(g.input.data), static_cast<Out*>(g.output.data), g.input.s[0],
                       g.input.s[1], g.input.s[2], g.input.s[3], g.output.s[0], g.output.s[1], g.output.s[2],
                       g.output.s[3], d, h, w, tiles_x, fs0, fs1, fs2, h0, h1, h2, g.post);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

} // namespace

int launch_remap(const RemapDesc& d, hipStream_t stream, std::string& err) {
    switch (elem_bytes(d.input.dtype)) {
        case 4: return run_remap<float>(d, stream, err);
        case 8: return run_remap<float2>(d, stream, err);
        case 16: return run_remap<double2>(d, stream, err);
        default: err = "remap: unsupported element type"; return NOA_HIP_ERROR_UNSUPPORTED;
    }
}

int launch_gridding_correction(const GriddingDesc& d, hipStream_t stream, std::string& err) {
    const bool i64_ = d.input.dtype == NOA_HIP_F64, o64 = d.output.dtype == NOA_HIP_F64;
    // coord_t = f64 if either side is f64 (FourierProject.hpp:2073).
    if (i64_ && o64) return run_gridding<double, double, double>(d, stream, err);
    if (i64_) return run_gridding<double, float, double>(d, stream, err);
    if (o64) return run_gridding<float, double, double>(d, stream, err);
    return run_gridding<float, float, float>(d, stream, err);
}

} // namespace noa_hip
