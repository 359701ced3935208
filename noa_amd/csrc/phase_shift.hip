// Phase shift of 2D rfft stacks: replaces noa::signal::phase_shift_2d (src/noa/signal/PhaseShift.hpp:215-289,
// operator :78-139) for the h2h, h2hc, hc2h and hc2hc layouts.
//
// One lane per element of the (B, H, W/2+1) stack, 256-lane workgroups over consecutive elements of one slice (x
// fastest): reads and writes are coalesced and the shift of the slice is uniform over the workgroup.
#include "common.hpp"
#include "kernels.hpp"

#include <cmath>

namespace noa_hip {

template<class V, class R>
struct PhaseShiftParams {
    const V* in; i64 is_b; i32 is_y, is_x;
    V* out; i64 os_b; i32 os_y, os_x;
    i32 h, w, wr, tiles;
    bool src_centered, dst_centered;
    const R* shifts; bool batched; R single[2];
    R norm_y, norm_x, cutoff_sqd;
    int mode;
};

// PS_EPT elements of one slice per lane (tiles of 256 * PS_EPT elements), the input loads issued before any phase
// is evaluated.
constexpr i32 PS_EPT = 4;

template<class V, class R>
__global__ __launch_bounds__(256) void phase_shift_kernel(const PhaseShiftParams<V, R> p) {
    const i32 b = static_cast<i32>(blockIdx.x) / p.tiles;
    const i32 first = (static_cast<i32>(blockIdx.x) - b * p.tiles) * (256 * PS_EPT) + static_cast<i32>(threadIdx.x);
    const i32 count = p.h * p.wr;
    V v[PS_EPT];
    i32 ys[PS_EPT], xs[PS_EPT];
#pragma unroll
    for (i32 k = 0; k < PS_EPT; ++k) {
        const i32 e = min(first + k * 256, count - 1);   // lanes past the end repeat the last element, not stored
        const i32 y = e / p.wr, x = e - y * p.wr;
        ys[k] = y;
        xs[k] = x;
        if (p.in) v[k] = p.in[static_cast<i64>(b) * p.is_b + y * p.is_y + x * p.is_x];
    }
    const R sy = p.batched ? p.shifts[2 * b] : p.single[0];
    const R sx = p.batched ? p.shifts[2 * b + 1] : p.single[1];
#pragma unroll
    for (i32 k = 0; k < PS_EPT; ++k) {
        if (first + k * 256 >= count) continue;
        const i32 y = ys[k], x = xs[k];
        const i32 fy = index2frequency(p.src_centered, y, p.h);
        const i32 oy = frequency2index(p.dst_centered, fy, p.h);
        V* o = p.out + static_cast<i64>(b) * p.os_b + oy * p.os_y + x * p.os_x;
        if (p.mode == 1) {
            if (p.in) *o = v[k];
            else { V one; one.x = 1; one.y = 0; *o = one; }
            continue;
        }
        const V phase = phase_shift_at<V, R>(p.mode, sy, sx, fy, x, p.norm_y, p.norm_x, p.cutoff_sqd);
        if (p.in) *o = cmul(v[k], phase);
        else *o = phase;
    }
}

namespace {

template<class V, class R>
int run(const PhaseShiftDesc& d, hipStream_t stream, std::string& err) {
    PhaseShiftParams<V, R> p{};
    p.in = static_cast<const V*>(d.input.data);
    p.is_b = d.input.s0; p.is_y = static_cast<i32>(d.input.s1); p.is_x = static_cast<i32>(d.input.s2);
    p.out = static_cast<V*>(d.output.data);
    p.os_b = d.output.s0; p.os_y = static_cast<i32>(d.output.s1); p.os_x = static_cast<i32>(d.output.s2);
    p.h = static_cast<i32>(d.h);
    p.w = static_cast<i32>(d.w);
    p.wr = p.w / 2 + 1;
    p.tiles = (p.h * p.wr + 256 * PS_EPT - 1) / (256 * PS_EPT);
    p.src_centered = d.src_centered;
    p.dst_centered = d.dst_centered;
    p.shifts = static_cast<const R*>(d.shift.data);
    p.batched = d.shift.batched;
    p.single[0] = static_cast<R>(d.shift.single[0]);
    p.single[1] = static_cast<R>(d.shift.single[1]);
    p.norm_y = R(1) / static_cast<R>(d.h);     // m_norm = 1 / shape (PhaseShift.hpp:108)
    p.norm_x = R(1) / static_cast<R>(d.w);
    const R cutoff = static_cast<R>(d.shift.cutoff);
    p.cutoff_sqd = cutoff * cutoff;
    p.mode = d.shift.mode;
    const i64 blocks = d.b * p.tiles;
    if (blocks == 0) return NOA_HIP_OK;
    hipLaunchKernelGGL((phase_shift_kernel<V, R>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

} // namespace

int launch_phase_shift(const PhaseShiftDesc& d, hipStream_t stream, std::string& err) {
    if (d.c64) return d.shift.f64 ? run<double2, double>(d, stream, err) : run<double2, float>(d, stream, err);
    return d.shift.f64 ? run<float2, double>(d, stream, err) : run<float2, float>(d, stream, err);
}

} // namespace noa_hip
