// Strided 4D copy of real f32/f64 arrays, used to stage a real FFT operand through a contiguous buffer when its
// layout is one that rocFFT's even-length real kernels read or write wrongly (odd outer strides or distance, or a
// non-unit x stride). One lane per element, x fastest.
#include "common.hpp"
#include "kernels.hpp"

namespace noa_hip {

template<class T>
__global__ __launch_bounds__(256) void strided_copy_kernel(const T* __restrict__ src, i64 s0, i64 s1, i64 s2, i64 s3,
                                                           T* __restrict__ dst, i64 d0, i64 d1, i64 d2, i64 d3,
                                                           i64 n1, i64 n2, i64 n3, i64 total) {
    const i64 i = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= total) return;
    const i64 x = i % n3, r = i / n3;
    const i64 y = r % n2, r2 = r / n2;
    const i64 z = r2 % n1, b = r2 / n1;
    dst[b * d0 + z * d1 + y * d2 + x * d3] = src[b * s0 + z * s1 + y * s2 + x * s3];
}

int launch_strided_copy(const void* src, const i64 ss[4], void* dst, const i64 ds[4], const i64 shape[4],
                        bool f64, hipStream_t stream, std::string& err) {
    const i64 total = shape[0] * shape[1] * shape[2] * shape[3];
    if (total == 0) return NOA_HIP_OK;
    const i64 blocks = (total + 255) / 256;
    if (blocks > 0xffffffffLL) { err = "fft staging copy: too many elements"; return NOA_HIP_ERROR_UNSUPPORTED; }
    if (f64)
        hipLaunchKernelGGL((strided_copy_kernel<double>), dim3(static_cast<u32>(blocks)), dim3(256), 0, stream,
                           static_cast<const double*>(src), ss[0], ss[1], ss[2], ss[3], static_cast<double*>(dst),
                           ds[0], ds[1], ds[2], ds[3], shape[1], shape[2], shape[3], total);
    else
        hipLaunchKernelGGL((strided_copy_kernel<float>), dim3(static_cast<u32>(blocks)), dim3(256), 0, stream,
                           static_cast<const float*>(src), ss[0], ss[1], ss[2], ss[3], static_cast<float*>(dst),
                           ds[0], ds[1], ds[2], ds[3], shape[1], shape[2], shape[3], total);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

} // namespace noa_hip
