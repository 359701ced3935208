// Fourier extraction of central slices from a 3D rfft volume (gather).
//
// Replaces the FourierExtract operator (src/noa/xform/FourierProject.hpp:541-725) launched by
// launch_extract_central_slices_3d (:1396-1487) through the CUDA iwise_{3,4}d_static kernels.
//
// Design (gfx950):
//   - One workgroup of 256 lanes (4 wave64) owns `ppt * 256` consecutive pixels of one slice (row-major
//     y, u), so stores are fully coalesced along u and a wavefront's gathers walk a line of the volume.
//   - The slice's rotation (and scaling) is uniform over the workgroup and is read with scalar loads.
//   - Workgroups are remapped so that each of the 8 XCDs processes a contiguous range of tiles: consecutive
//     tiles of a slice share one L2, which holds the volume neighbourhood of that slice (the whole volume
//     stays resident in the 256 MB Infinity Cache).
//   - The w-windowed-sinc variant (window size > 1, :661-694) loops over the window inside the lane that owns
//     the pixel, in increasing w, so no zero fill and no atomics are needed; the sum order is that of noa's
//     serial CPU loop.
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace noa_hip {

template<class C>
struct ExtractParams {
    const void* vol; i32 vs0, vs1, vs2;
    const void* wvol; i32 wvs0, wvs1, wvs2;
    void* out; i64 os_b; i32 os_y, os_x;
    void* wout; i64 ws_b; i32 ws_y, ws_x;
    i32 grid[3];
    C target[3];
    C slice_f[2];
    i32 slice_h, slice_wr, pixels, tiles, ppt;
    float inv_wr;          // 1 / slice_wr (lean kernel's pixel -> row)
    C cutoff_sqd;
    C ews[2];
    C fftfreq_sinc, fftfreq_blackman;
    i32 window;            // blackman window size along w (1 = plain 3D extraction)
    int interp_mode;       // the method of the INTERP_WINDOW instantiation (CUBIC, CUBIC_BSPLINE, LANCZOS*)
    XformArg<C> rot, scale;
    bool general, vol_centered, slice_centered, has_wvol, has_wout;
    // Fused phase shift of the output slices (phase_shift_2d in the slices' own layout); shift_mode < 0: none.
    int shift_mode;
    bool shift_f64, shift_batched;
    const void* shifts;
    float shift_f[2], norm_f[2], cutoff_sqd_f;
    double shift_d[2], norm_d[2], cutoff_sqd_d;
};

// The phase shift of slice b at pixel (fy, u) applied to an extracted complex value (PhaseShift.hpp:113-131).
template<class C, class Out>
__device__ __forceinline__ Out shift_value(const ExtractParams<C>& p, i32 b, i32 fy, i32 u, Out value) {
    if constexpr (vt<Out>::cx) {
        if (p.shift_mode < 0 || p.shift_mode == 1) return value;
        Out phase;
        if (p.shift_f64) {
            const double* s = static_cast<const double*>(p.shifts);
            const double sy = p.shift_batched ? s[2 * b] : p.shift_d[0], sx = p.shift_batched ? s[2 * b + 1] : p.shift_d[1];
            phase = phase_shift_at<Out, double>(p.shift_mode, sy, sx, fy, u, p.norm_d[0], p.norm_d[1], p.cutoff_sqd_d);
        } else {
            const float* s = static_cast<const float*>(p.shifts);
            const float sy = p.shift_batched ? s[2 * b] : p.shift_f[0], sx = p.shift_batched ? s[2 * b + 1] : p.shift_f[1];
            phase = phase_shift_at<Out, float>(p.shift_mode, sy, sx, fy, u, p.norm_f[0], p.norm_f[1], p.cutoff_sqd_f);
        }
        return cmul(value, phase);
    } else {
        (void)p; (void)b; (void)fy; (void)u;
        return value;
    }
}

template<class C, class S, class M, class Out, class W, int INTERP>
__global__ __launch_bounds__(256) void extract_kernel(const ExtractParams<C> p) {
    const u32 lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const i32 b = static_cast<i32>(lb / static_cast<u32>(p.tiles));
    const i32 tile = static_cast<i32>(lb) - b * p.tiles;

    const Rot<C> rot = load_rotation(p.rot, b);
    Scale<C> sc;
    sc.identity = true;
    if (p.general) sc = load_scaling(p.scale, b);
    const Bounds3 bd = make_bounds3(p.grid[0], p.grid[1], p.grid[2]);

    const S* __restrict__ vol = static_cast<const S*>(p.vol);
    const W* __restrict__ wvol = static_cast<const W*>(p.wvol);
    Out* __restrict__ out = static_cast<Out*>(p.out) + static_cast<i64>(b) * p.os_b;
    W* __restrict__ wout = p.has_wout ? static_cast<W*>(p.wout) + static_cast<i64>(b) * p.ws_b : nullptr;

    const i32 first = tile * p.ppt * 256 + static_cast<i32>(threadIdx.x);
    if (p.window == 1) {
        // 3D operator (:632-657): overwrite, 0 beyond the cutoff. The tap loads of the lane's pixels (up to
        // MAXP) are all issued before the first is used.
        constexpr i32 MAXP = 4;
        Taps3<M, C> tv[MAXP];
        Taps3<W, C> tw[MAXP];
        bool inside[MAXP];
        i32 ys[MAXP], us[MAXP];
        if constexpr (INTERP == INTERP_WINDOW) {
            // Windows of 4 to 8 taps per dimension: one pixel at a time.
            for (i32 k = 0; k < p.ppt; ++k) {
                const i32 pix = first + k * 256;
                if (pix >= p.pixels) break;
                const i32 y = pix / p.slice_wr, u = pix - y * p.slice_wr;
                vec2<C> f2;
                f2.y = static_cast<C>(index2frequency(p.slice_centered, y, p.slice_h)) / p.slice_f[0];
                f2.x = static_cast<C>(u) / p.slice_f[1];
                const vec3<C> f = p.general ? slice2grid<true>(f2, sc, rot, p.ews[0], p.ews[1])
                                            : slice2grid<false>(f2, sc, rot, p.ews[0], p.ews[1]);
                Out value = vzero<Out>();
                W weight{};
                if (dot3(f) <= p.cutoff_sqd) {
                    const C fz = f.z * p.target[0], fy = f.y * p.target[1], fx = f.x * p.target[2];
                    value = vcast<Out>(interp3<INTERP, M>(vol, p.vs0, p.vs1, p.vs2, bd, p.vol_centered, fz, fy, fx,
                                                          p.interp_mode));
                    if (p.has_wout)
                        weight = p.has_wvol ? interp3<INTERP, W>(wvol, p.wvs0, p.wvs1, p.wvs2, bd, p.vol_centered, fz,
                                                                 fy, fx, p.interp_mode)
                                            : W(1);
                }
                if (p.shift_mode >= 0)
                    value = shift_value(p, b, index2frequency(p.slice_centered, y, p.slice_h), u, value);
                out[y * p.os_y + u * p.os_x] = value;
                if (p.has_wout) wout[y * p.ws_y + u * p.ws_x] = weight;
            }
            return;
        }
#pragma unroll
        for (i32 k = 0; k < MAXP; ++k) {
            const i32 pix = first + k * 256;
            const bool live = k < p.ppt && pix < p.pixels;
            const i32 y = live ? pix / p.slice_wr : 0;
            const i32 u = live ? pix - y * p.slice_wr : 0;
            ys[k] = live ? y : -1;
            us[k] = u;
            // compute_fftfreq_in_volume_ (:699-705)
            vec2<C> f2;
            f2.y = static_cast<C>(index2frequency(p.slice_centered, y, p.slice_h)) / p.slice_f[0];
            f2.x = static_cast<C>(u) / p.slice_f[1];
            const vec3<C> f = p.general ? slice2grid<true>(f2, sc, rot, p.ews[0], p.ews[1])
                                        : slice2grid<false>(f2, sc, rot, p.ews[0], p.ews[1]);
            inside[k] = live && dot3(f) <= p.cutoff_sqd;
            const C fz = f.z * p.target[0], fy = f.y * p.target[1], fx = f.x * p.target[2];
            tv[k] = fetch3<INTERP, M>(GlobalLoad<S>{vol}, p.vs0, p.vs1, p.vs2, bd, p.vol_centered, fz, fy, fx,
                                      inside[k]);
            if (p.has_wout && p.has_wvol)
                tw[k] = fetch3<INTERP, W>(GlobalLoad<W>{wvol}, p.wvs0, p.wvs1, p.wvs2, bd, p.vol_centered, fz, fy, fx,
                                          inside[k]);
        }
#pragma unroll
        for (i32 k = 0; k < MAXP; ++k) {
            if (ys[k] < 0) continue;
            Out value = inside[k] ? vcast<Out>(combine3<INTERP>(tv[k])) : vzero<Out>();
            if (p.shift_mode >= 0)
                value = shift_value(p, b, index2frequency(p.slice_centered, ys[k], p.slice_h), us[k], value);
            out[ys[k] * p.os_y + us[k] * p.os_x] = value;
            if (p.has_wout) {
                W weight{};
                if (inside[k]) weight = p.has_wvol ? combine3<INTERP>(tw[k]) : W(1);
                wout[ys[k] * p.ws_y + us[k] * p.ws_x] = weight;
            }
        }
        return;
    }
    for (i32 k = 0; k < p.ppt; ++k) {
        const i32 pix = first + k * 256;
        if (pix >= p.pixels) break;
        const i32 y = pix / p.slice_wr;
        const i32 u = pix - y * p.slice_wr;

        // compute_fftfreq_in_volume_ (:699-705)
        vec2<C> f2;
        f2.y = static_cast<C>(index2frequency(p.slice_centered, y, p.slice_h)) / p.slice_f[0];
        f2.x = static_cast<C>(u) / p.slice_f[1];
        const vec3<C> f = p.general ? slice2grid<true>(f2, sc, rot, p.ews[0], p.ews[1])
                                    : slice2grid<false>(f2, sc, rot, p.ews[0], p.ews[1]);

        // 4D operator (:661-694), summed over w in increasing order.
        using RO = typename vt<Out>::real;
        Out value = vzero<Out>();
        W weight{};
        for (i32 w = 0; w < p.window; ++w) {
            const C offset = static_cast<C>(w - p.window / 2) / p.target[0];   // w_index_to_fftfreq_offset
            vec3<C> fw = f;
            fw.z += offset;
            if (dot3(fw) > p.cutoff_sqd) continue;
            const C fz = fw.z * p.target[0], fy = fw.y * p.target[1], fx = fw.x * p.target[2];
            const C conv = windowed_sinc<C>(offset, p.fftfreq_sinc, p.fftfreq_blackman);
            const Out v = vcast<Out>(interp3<INTERP, M>(vol, p.vs0, p.vs1, p.vs2, bd, p.vol_centered, fz, fy, fx,
                                                        p.interp_mode));
            value = vadd(value, vmul(v, static_cast<RO>(conv)));
            if (p.has_wout) {
                const W wv = p.has_wvol ? interp3<INTERP, W>(wvol, p.wvs0, p.wvs1, p.wvs2, bd, p.vol_centered, fz, fy, fx,
                                                             p.interp_mode)
                                        : W(1);
                weight = weight + wv * static_cast<W>(conv);
            }
        }
        if (p.shift_mode >= 0) value = shift_value(p, b, index2frequency(p.slice_centered, y, p.slice_h), u, value);
        out[y * p.os_y + u * p.os_x] = value;
        if (p.has_wout) wout[y * p.ws_y + u * p.ws_x] = weight;
    }
}

// The common configuration on a lean path: complex FP32 volume and slices, no weights, no scaling or EWS, no w
// window, LINEAR, unit x stride; the fused phase shift in FP32 (SHIFT 1: per-slice or broadcast shifts, SHIFT 2:
// the half-shape +-1 pattern, SHIFT 0: none). Same expressions as extract_kernel (bit-identical results), minus the
// work that the general kernel does for every pixel: the integer division of the pixel index (here a float
// reciprocal, exact for slices below 2^21 pixels, checked on the host), the runtime selection between the plain and
// the scaled/EWS transforms, the weight paths, and the runtime choice of the shift precision.
// Copy of a complex volume (unit x stride; c32 or a pair of 16-bit floats) into the block layout of BrickLayout:
// one lane per element of the stored (nz, ny, wr) volume; the block padding (sizes not multiples of the block) is
// never read.
template<class S>
__global__ __launch_bounds__(256) void brick_volume_kernel(const S* __restrict__ in, i32 vs0, i32 vs1,
                                                           S* __restrict__ out, i32 ny, i32 wr, i64 total,
                                                           BrickLayout L) {
    const i64 i = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= total) return;
    const i32 x = static_cast<i32>(i % wr);
    const i64 r = i / wr;
    const i32 y = static_cast<i32>(r % ny), z = static_cast<i32>(r / ny);
    out[L.z(z) + L.y(y) + L.x(x)] = in[static_cast<i64>(z) * vs0 + static_cast<i64>(y) * vs1 + x];
}

template<int SHIFT, bool BRICK, class S>
__global__ __launch_bounds__(256) void extract_c32_kernel(const ExtractParams<float> p) {
    const u32 lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const i32 b = static_cast<i32>(lb / static_cast<u32>(p.tiles));
    const i32 tile = static_cast<i32>(lb) - b * p.tiles;

    const Rot<float> rot = load_rotation(p.rot, b);
    Scale<float> sc;
    sc.identity = true;
    const Bounds3 bd = make_bounds3(p.grid[0], p.grid[1], p.grid[2]);
    float sy = 0, sx = 0;
    if constexpr (SHIFT == 1) {
        const float* s = static_cast<const float*>(p.shifts);
        sy = p.shift_batched ? s[2 * b] : p.shift_f[0];
        sx = p.shift_batched ? s[2 * b + 1] : p.shift_f[1];
    }
    const S* __restrict__ vol = static_cast<const S*>(p.vol);
    float2* __restrict__ out = static_cast<float2*>(p.out) + static_cast<i64>(b) * p.os_b;

    constexpr i32 MAXP = 4;
    Taps3<float2, float> tv[MAXP];
    bool inside[MAXP];
    i32 ys[MAXP], us[MAXP];
    const i32 first = tile * p.ppt * 256 + static_cast<i32>(threadIdx.x);
#pragma unroll
    for (i32 k = 0; k < MAXP; ++k) {
        const i32 pix = first + k * 256;
        const bool live = k < p.ppt && pix < p.pixels;
        // pix / wr from a float reciprocal: (pix + 0.5) / wr is at least 0.5 / wr away from an integer, far above
        // the rounding error for pix < 2^21.
        const i32 yq = static_cast<i32>((static_cast<float>(pix) + 0.5f) * p.inv_wr);
        const i32 y = live ? yq : 0;
        const i32 u = live ? pix - yq * p.slice_wr : 0;
        ys[k] = live ? y : -1;
        us[k] = u;
        vec2<float> f2;
        const i32 fy = index2frequency(p.slice_centered, y, p.slice_h);
        f2.y = static_cast<float>(fy) / p.slice_f[0];
        f2.x = static_cast<float>(u) / p.slice_f[1];
        const vec3<float> f = slice2grid<false>(f2, sc, rot, 0.f, 0.f);
        inside[k] = live && dot3(f) <= p.cutoff_sqd;
        const float fz = f.z * p.target[0], fyv = f.y * p.target[1], fx = f.x * p.target[2];
        if constexpr (BRICK)
            tv[k] = fetch3<NOA_HIP_INTERP_LINEAR, float2>(GlobalLoad<S>{vol}, BrickLayout{p.vs0, p.vs1}, bd,
                                                          p.vol_centered, fz, fyv, fx, inside[k]);
        else
            tv[k] = fetch3<NOA_HIP_INTERP_LINEAR, float2>(GlobalLoad<S>{vol}, p.vs0, p.vs1, 1, bd, p.vol_centered,
                                                          fz, fyv, fx, inside[k]);
    }
#pragma unroll
    for (i32 k = 0; k < MAXP; ++k) {
        if (ys[k] < 0) continue;
        float2 value = inside[k] ? combine3<NOA_HIP_INTERP_LINEAR>(tv[k]) : vzero<float2>();
        if constexpr (SHIFT != 0) {
            const i32 fy = index2frequency(p.slice_centered, ys[k], p.slice_h);
            value = cmul(value, phase_shift_at<float2, float>(SHIFT == 1 ? 0 : 2, sy, sx, fy, us[k], p.norm_f[0],
                                                               p.norm_f[1], p.cutoff_sqd_f));
        }
        out[ys[k] * p.os_y + us[k] * p.os_x] = value;
    }
}

namespace {

template<class C>
XformArg<C> to_arg(const Xform& x) {
    XformArg<C> a{};
    a.data = x.data;
    a.kind = x.kind;
    a.batched = x.batched ? 1 : 0;
    for (int i = 0; i < 9; ++i) a.single[i] = x.single[i];
    return a;
}

template<class C, class S, class M, class Out, class W>
int run(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    ExtractParams<C> p{};
    p.vol = d.input.data; p.vs0 = (i32)d.input.s0; p.vs1 = (i32)d.input.s1; p.vs2 = (i32)d.input.s2;
    p.has_wvol = d.input_weight.data != nullptr;
    p.wvol = d.input_weight.data; p.wvs0 = (i32)d.input_weight.s0; p.wvs1 = (i32)d.input_weight.s1; p.wvs2 = (i32)d.input_weight.s2;
    p.out = d.output.data; p.os_b = d.output.s0; p.os_y = (i32)d.output.s1; p.os_x = (i32)d.output.s2;
    p.has_wout = d.output_weight.data != nullptr;
    p.wout = d.output_weight.data; p.ws_b = d.output_weight.s0; p.ws_y = (i32)d.output_weight.s1; p.ws_x = (i32)d.output_weight.s2;
    for (int i = 0; i < 3; ++i) {
        p.grid[i] = static_cast<i32>(d.in_shape[i + 1]);
        p.target[i] = static_cast<C>(d.target[i]);
    }
    p.slice_f[0] = static_cast<C>(d.out_shape[2]);
    p.slice_f[1] = static_cast<C>(d.out_shape[3]);
    p.slice_h = static_cast<i32>(d.out_shape[2]);
    p.slice_wr = static_cast<i32>(d.out_shape[3] / 2 + 1);
    p.pixels = p.slice_h * p.slice_wr;
    p.shift_mode = d.shift.present ? d.shift.mode : -1;
    p.shift_f64 = d.shift.f64;
    p.shift_batched = d.shift.batched;
    p.shifts = d.shift.data;
    for (int i = 0; i < 2; ++i) {
        p.shift_f[i] = static_cast<float>(d.shift.single[i]);
        p.shift_d[i] = d.shift.single[i];
    }
    // m_norm = 1 / shape in the shift precision (PhaseShift.hpp:108).
    p.norm_f[0] = 1.f / static_cast<float>(d.out_shape[2]);
    p.norm_f[1] = 1.f / static_cast<float>(d.out_shape[3]);
    p.norm_d[0] = 1.0 / static_cast<double>(d.out_shape[2]);
    p.norm_d[1] = 1.0 / static_cast<double>(d.out_shape[3]);
    p.cutoff_sqd_f = static_cast<float>(d.shift.cutoff) * static_cast<float>(d.shift.cutoff);
    p.cutoff_sqd_d = d.shift.cutoff * d.shift.cutoff;
    C cutoff = std::max(static_cast<C>(d.cutoff), C(0));
    p.cutoff_sqd = cutoff * cutoff;
    p.general = d.general;
    if (d.general && (d.ews_radius[0] != 0 || d.ews_radius[1] != 0)) {
        p.ews[0] = 1 / (2 * static_cast<C>(d.ews_radius[0]));
        p.ews[1] = 1 / (2 * static_cast<C>(d.ews_radius[1]));
    }
    // Along the w of the grid (:622-626).
    p.fftfreq_sinc = std::max(static_cast<C>(d.fftfreq_sinc), 1 / p.target[0]);
    p.fftfreq_blackman = std::max(static_cast<C>(d.fftfreq_blackman), 1 / p.target[0]);
    p.window = static_cast<i32>(blackman_window_size(static_cast<double>(p.fftfreq_blackman),
                                                     static_cast<double>(p.target[0])));
    p.rot = to_arg<C>(d.rot);
    p.scale = to_arg<C>(d.scale);
    p.vol_centered = d.src_centered;
    p.slice_centered = d.dst_centered;

    // Pixels per lane: the largest of {4, 3, 2, 1} that tiles a slice without a partial workgroup, else 4.
    p.ppt = 4;
    for (int c : {4, 3, 2, 1}) {
        if (p.pixels % (256 * c) == 0) { p.ppt = c; break; }
    }
    p.tiles = (p.pixels + 256 * p.ppt - 1) / (256 * p.ppt);
    const i64 n_slices = d.out_shape[0];
    const i64 blocks = n_slices * p.tiles;
    if (blocks == 0) return NOA_HIP_OK;
    if (blocks > 0x7fffffffLL) { err = "too many slices for one launch"; return NOA_HIP_ERROR_UNSUPPORTED; }

    const dim3 grid(static_cast<unsigned>(blocks)), block(256);
    // c32 volumes and the 16-bit pair storages (FP32 math after the load, as in the general kernel).
    constexpr bool LEAN_TYPES = std::is_same_v<C, float> && std::is_same_v<Out, float2> && std::is_same_v<M, float2> &&
                                (std::is_same_v<S, float2> || std::is_same_v<S, __half2> ||
                                 std::is_same_v<S, __hip_bfloat162>);
    if constexpr (LEAN_TYPES) {
        const bool lean = d.interp == NOA_HIP_INTERP_LINEAR && !p.general && p.window == 1 && !p.has_wvol &&
                          !p.has_wout && p.vs2 == 1 && p.pixels < (1 << 21) &&
                          (p.shift_mode < 0 || p.shift_mode == 1 || p.shift_mode == 2 || !p.shift_f64);
        if (lean) {
            p.inv_wr = 1.f / static_cast<float>(p.slice_wr);
            // Many slices from one volume: gather from a block copy of it (BrickLayout) (one read and one write of the
            // volume, then every trilinear gather touches fewer cache lines). Worth it once the slices hold about
            // twice the volume's elements; NOA_HIP_EXTRACT_LAYOUT=linear|bricks forces either.
            const i32 nz = p.grid[0], ny = p.grid[1], wr = p.grid[2] / 2 + 1;
            const i64 stored = static_cast<i64>(nz) * ny * wr;
            i64 bz, by, brick_elems;
            BrickLayout::strides(nz, ny, wr, bz, by, brick_elems);
            const char* env = std::getenv("NOA_HIP_EXTRACT_LAYOUT");
            bool bricks = n_slices * static_cast<i64>(p.pixels) >= 2 * stored;
            if (env && std::strcmp(env, "linear") == 0) bricks = false;
            if (env && std::strcmp(env, "bricks") == 0) bricks = true;
            bricks = bricks && brick_elems < 0x7fffffffLL;
            void* brick = nullptr;
            if (bricks) {
                if (hipMallocAsync(&brick, static_cast<size_t>(brick_elems) * sizeof(S), stream) != hipSuccess) {
                    (void)hipGetLastError();
                    brick = nullptr;   // not enough memory for the copy: gather from the row-major volume
                }
            }
            if (brick) {
                const BrickLayout L{static_cast<i32>(bz), static_cast<i32>(by)};
                hipLaunchKernelGGL((brick_volume_kernel<S>), dim3(static_cast<u32>((stored + 255) / 256)), dim3(256), 0,
                                   stream, static_cast<const S*>(p.vol), p.vs0, p.vs1, static_cast<S*>(brick),
                                   ny, wr, stored, L);
                ExtractParams<C> q = p;
                q.vol = brick;
                q.vs0 = L.bz;
                q.vs1 = L.by;
                if (q.shift_mode == 0) hipLaunchKernelGGL((extract_c32_kernel<1, true, S>), grid, block, 0, stream, q);
                else if (q.shift_mode == 2) hipLaunchKernelGGL((extract_c32_kernel<2, true, S>), grid, block, 0, stream, q);
                else hipLaunchKernelGGL((extract_c32_kernel<0, true, S>), grid, block, 0, stream, q);
                (void)hipFreeAsync(brick, stream);
            } else {
                if (p.shift_mode == 0) hipLaunchKernelGGL((extract_c32_kernel<1, false, S>), grid, block, 0, stream, p);
                else if (p.shift_mode == 2) hipLaunchKernelGGL((extract_c32_kernel<2, false, S>), grid, block, 0, stream, p);
                else hipLaunchKernelGGL((extract_c32_kernel<0, false, S>), grid, block, 0, stream, p);
            }
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
            return NOA_HIP_OK;
        }
    }
    p.interp_mode = d.interp;
    if (d.interp == NOA_HIP_INTERP_NEAREST)
        hipLaunchKernelGGL((extract_kernel<C, S, M, Out, W, NOA_HIP_INTERP_NEAREST>), grid, block, 0, stream, p);
    else if (d.interp == NOA_HIP_INTERP_LINEAR)
        hipLaunchKernelGGL((extract_kernel<C, S, M, Out, W, NOA_HIP_INTERP_LINEAR>), grid, block, 0, stream, p);
    else
        hipLaunchKernelGGL((extract_kernel<C, S, M, Out, W, INTERP_WINDOW>), grid, block, 0, stream, p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class C>
int dispatch(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    const int in = d.input.dtype, out = d.output.dtype;
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return run<C, float2, float2, float2, float>(d, stream, err);
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return run<C, float2, float2, float, float>(d, stream, err);
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return run<C, float, float, float, float>(d, stream, err);
    if (in == NOA_HIP_C16 && out == NOA_HIP_C32) return run<C, __half2, float2, float2, float>(d, stream, err);
    if (in == NOA_HIP_CBF16 && out == NOA_HIP_C32) return run<C, __hip_bfloat162, float2, float2, float>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return run<C, double2, double2, double2, double>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return run<C, double2, double2, double, double>(d, stream, err);
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return run<C, double, double, double, double>(d, stream, err);
    err = "extract: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_extract(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    return d.coord_f64 ? dispatch<double>(d, stream, err) : dispatch<float>(d, stream, err);
}

} // namespace noa_hip
