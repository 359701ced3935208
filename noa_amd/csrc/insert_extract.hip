// Extraction of central slices from a virtual volume made of other central slices: noa's
// insert_and_extract_central_slices_3d (src/noa/xform/FourierProject.hpp:1916-2049), launcher :1489-1598,
// operator FourierInsertExtract :727-1006.
//
// noa evaluates, for every output pixel, every input slice (O(pixels x slices)). Here one wave owns an 8x8 tile
// of output pixels (8 consecutive frequencies along y, 8 along x) and first scans the input slices 64 at a time,
// one slice per lane: a slice is a candidate when the window along its normal, |fz| < fftfreq_blackman, can hold
// a pixel of the tile. Without an Ewald sphere fz is a linear function of the output pixel's frequency, so over
// the tile its extremes are at the four corners (at each w offset); a margin far above the rounding
// differences between the corners and the interior keeps the test conservative. The candidates are then
// evaluated in increasing slice order, the rotation broadcast from the lane that scanned it, with noa's
// expressions and noa's exact test per pixel, so the sums run in noa's order and skipping a slice never changes a
// result. With an Ewald sphere every slice is a candidate.
//
// The w-windowed (4D) form keeps the window sum in registers: the outputs start from 0 (or from their value when
// adding) and the w terms are added in increasing w, the order of noa's serial CPU path; no atomics are needed
// since one lane owns each output pixel.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace noa_hip {

constexpr int IE_TILE = 8;                 // 8 x 8 output pixels per wave
constexpr int IE_WAVES = 4;                // waves per workgroup

template<class C, class In, class W>
struct InsExtParams {
    // Input slices (batch, y, x strides; the batch stride is 0 for a single slice, as to_interpolator_input).
    const In* in; i64 is_b; i32 is_y, is_x; bool in_is_value; In in_value;
    const W* win; i64 wis_b; i32 wis_y, wis_x; bool win_present, win_is_value; W win_value;
    i32 in_h, in_w;                        // logical height and width of the input slices
    C in_f[2];
    i32 n_in;
    bool in_centered;
    // Output slices.
    void* out; i64 os_b; i32 os_y, os_x;
    W* wout; i64 wos_b; i32 wos_y, wos_x;
    i32 out_h, out_wr;                     // logical height, rfft width
    C out_f[2];
    bool out_centered;
    i32 n_out, tiles_y, tiles_x;
    // Transforms.
    XformArg<C> rot_in, scale_in, rot_out, scale_out;
    C ews[2];
    C cutoff_sqd, in_sinc, in_blackman, w_sinc, w_blackman, volume_z;
    i32 w_size;
    int interp_mode;                       // the method of the INTERP_WINDOW instantiation
    bool add, correct, cull;
};

template<class C, class In, class Out, class W, int INTERP, bool GENERAL>
__global__ __launch_bounds__(IE_TILE * IE_TILE * IE_WAVES) void insert_extract_kernel(const InsExtParams<C, In, W> p) {
    using RIn = typename vt<In>::real;
    using ROut = typename vt<Out>::real;
    const i32 lane = static_cast<i32>(threadIdx.x & 63);
    const i64 tile = static_cast<i64>(blockIdx.x) * IE_WAVES + (threadIdx.x >> 6);
    const i64 tiles_per_slice = static_cast<i64>(p.tiles_y) * p.tiles_x;
    if (tile >= tiles_per_slice * p.n_out) return;   // whole waves; the kernel has no barrier
    const i32 b = static_cast<i32>(tile / tiles_per_slice);
    const i32 t = static_cast<i32>(tile - static_cast<i64>(b) * tiles_per_slice);
    const i32 ty = t / p.tiles_x, tx = t - (t / p.tiles_x) * p.tiles_x;
    const i32 fy0 = -(p.out_h / 2) + ty * IE_TILE, x0 = tx * IE_TILE;
    const i32 fy = fy0 + (lane >> 3), x = x0 + (lane & 7);
    const bool valid = fy <= (p.out_h - 1) / 2 && x < p.out_wr;

    // The output pixel's frequency in the virtual volume (compute_fftfreq_in_volume_, :927-933).
    Scale<C> so;
    so.identity = true;
    if constexpr (GENERAL) so = load_scaling(p.scale_out, b);
    const Rot<C> ro = load_rotation(p.rot_out, b);
    auto to_grid = [&](i32 frequency_y, i32 frequency_x) {
        vec2<C> f2;
        f2.y = static_cast<C>(frequency_y) / p.out_f[0];
        f2.x = static_cast<C>(frequency_x) / p.out_f[1];
        return slice2grid<GENERAL>(f2, so, ro, p.ews[0], p.ews[1]);
    };
    const vec3<C> f3 = to_grid(fy, x);
    vec3<C> corner[4];
    corner[0] = to_grid(fy0, x0);
    corner[1] = to_grid(fy0, x0 + IE_TILE - 1);
    corner[2] = to_grid(fy0 + IE_TILE - 1, x0);
    corner[3] = to_grid(fy0 + IE_TILE - 1, x0 + IE_TILE - 1);

    const i32 oy = frequency2index(p.out_centered, fy, p.out_h);
    Out* const out = static_cast<Out*>(p.out) + (static_cast<i64>(b) * p.os_b + oy * p.os_y + x * p.os_x);
    W* const wout = p.wout ? p.wout + (static_cast<i64>(b) * p.wos_b + oy * p.wos_y + x * p.wos_x) : nullptr;
    const bool four_d = p.w_size > 1;
    const bool quat = p.rot_in.kind == NOA_HIP_QUAT_F32 || p.rot_in.kind == NOA_HIP_QUAT_F64;
    Out acc = vzero<Out>();
    W wacc{};
    if (four_d && p.add && valid) {
        acc = *out;
        if (wout) wacc = *wout;
    }

    for (i32 w = 0; w < p.w_size; ++w) {
        vec3<C> f = f3;
        C offset = 0;
        if (four_d) {
            offset = static_cast<C>(w - p.w_size / 2) / p.volume_z;   // w_index_to_fftfreq_offset (:142-145)
            f.z += offset;
        }
        const bool inside = valid && dot3(f) <= p.cutoff_sqd;

        // Bound of fz over the tile at this offset: a slice whose interval misses (-fb, fb) is not a candidate.
        C mag = 0;
        vec3<C> cw[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            cw[c] = corner[c];
            cw[c].z += offset;
            mag = fmax(mag, fabs(cw[c].z) + fabs(cw[c].y) + fabs(cw[c].x));
        }
        const C margin = static_cast<C>(1e-5) * (1 + mag);

        // Every lane scans (slice base + lane), so the scan runs with the whole wave, inside the cutoff or not.
        In value = vzero<In>();
        W weight{};
        const i32 n_scan = __any(inside) ? p.n_in : 0;
        for (i32 base = 0; base < n_scan; base += 64) {
            const i32 i = base + lane;
            Rot<C> r{};
            bool cand = false;
            if (i < p.n_in) {
                r = load_rotation(p.rot_in, i);
                cand = true;
                if (p.cull) {
                    C lo = r.apply(cw[0]).z, hi = lo;
#pragma unroll
                    for (int c = 1; c < 4; ++c) {
                        const C z = r.apply(cw[c]).z;
                        lo = fmin(lo, z);
                        hi = fmax(hi, z);
                    }
                    cand = lo - margin < p.in_blackman && hi + margin > -p.in_blackman;
                }
            }
            u64 mask = __ballot(cand);
            while (mask) {
                const i32 k = __ffsll(static_cast<long long>(mask)) - 1;
                mask &= mask - 1;
                const i32 ii = base + k;
                Rot<C> rk;
                rk.quat = quat;
#pragma unroll
                for (int q = 0; q < 9; ++q) rk.m[q] = bcast(r.m[q], k);
                Scale<C> si;
                si.identity = true;
                if constexpr (GENERAL) si = load_scaling(p.scale_in, ii);
                // sample_virtual_volume_ (:935-977).
                vec2<C> g;
                const C fz = grid2slice<GENERAL>(f, si, rk, p.ews[0], p.ews[1], g);
                if (inside && fabs(fz) < p.in_blackman) {
                    const C window = windowed_sinc<C>(fz, p.in_sinc, p.in_blackman);
                    const C py = g.y * p.in_f[0], px = g.x * p.in_f[1];
                    const In v = interp2<INTERP, In>(p.in + static_cast<i64>(ii) * p.is_b, p.is_y, p.is_x, p.in_h,
                                                     p.in_w, p.in_centered, py, px, p.in_is_value, p.in_value,
                                                     p.interp_mode);
                    value = vadd(value, vmul(v, static_cast<RIn>(window)));
                    if (p.win_present) {
                        const W wv = interp2<INTERP, W>(p.win + static_cast<i64>(ii) * p.wis_b, p.wis_y, p.wis_x,
                                                        p.in_h, p.in_w, p.in_centered, py, px, p.win_is_value,
                                                        p.win_value, p.interp_mode);
                        weight = weight + wv * static_cast<W>(window);
                    } else {
                        weight = weight + static_cast<W>(window);
                    }
                }
            }
        }

        if (!four_d) {
            // 3D operator (:858-886): 0 beyond the cutoff, unless adding.
            if (!valid) continue;
            if (!inside) {
                if (!p.add) {
                    *out = vzero<Out>();
                    if (wout) *wout = W{};
                }
                continue;
            }
            if (p.correct) {
                const W final_weight = fmax(W(1), fabs(weight));
                if constexpr (vt<In>::cx) {
                    value.x = value.x / static_cast<RIn>(final_weight);
                    value.y = value.y / static_cast<RIn>(final_weight);
                } else {
                    value = value / static_cast<RIn>(final_weight);
                }
            }
            const Out o = vcast<Out>(value);
            *out = p.add ? vadd(*out, o) : o;
            if (wout) *wout = p.add ? *wout + weight : weight;
        } else if (inside) {
            // 4D operator (:890-924): the w term, weighted by the w window.
            const C conv = windowed_sinc<C>(offset, p.w_sinc, p.w_blackman);
            acc = vadd(acc, vmul(vcast<Out>(value), static_cast<ROut>(conv)));
            wacc = wacc + weight * static_cast<W>(conv);
        }
    }
    if (four_d && valid) {
        *out = acc;
        if (wout) *wout = wacc;
    }
}

namespace {

template<class C>
XformArg<C> xform_arg(const Xform& x) {
    XformArg<C> a{};
    a.data = x.data;
    a.kind = x.kind;
    a.batched = x.batched ? 1 : 0;
    for (int i = 0; i < 9; ++i) a.single[i] = x.single[i];
    return a;
}

template<class T> T constant(const Arr& a) {
    if constexpr (vt<T>::cx) {
        T v;
        v.x = static_cast<typename vt<T>::real>(a.value[0]);
        v.y = static_cast<typename vt<T>::real>(a.value[1]);
        return v;
    } else {
        return static_cast<T>(a.value[0]);
    }
}

template<class C, class In, class Out, class W>
int run(const InsertExtractDesc& d, hipStream_t stream, std::string& err) {
    InsExtParams<C, In, W> p{};
    p.in = static_cast<const In*>(d.input.data);
    p.is_b = d.in_shape[0] == 1 ? 0 : d.input.s0;
    p.is_y = static_cast<i32>(d.input.s1);
    p.is_x = static_cast<i32>(d.input.s2);
    p.in_is_value = d.input.is_value;
    p.in_value = constant<In>(d.input);
    p.win_present = d.input_weight.present();
    p.win = static_cast<const W*>(d.input_weight.data);
    p.wis_b = d.in_shape[0] == 1 ? 0 : d.input_weight.s0;
    p.wis_y = static_cast<i32>(d.input_weight.s1);
    p.wis_x = static_cast<i32>(d.input_weight.s2);
    p.win_is_value = d.input_weight.is_value;
    p.win_value = constant<W>(d.input_weight);
    p.in_h = static_cast<i32>(d.in_shape[2]);
    p.in_w = static_cast<i32>(d.in_shape[3]);
    p.in_f[0] = static_cast<C>(d.in_shape[2]);
    p.in_f[1] = static_cast<C>(d.in_shape[3]);
    p.n_in = static_cast<i32>(d.in_shape[0]);
    p.in_centered = d.src_centered;
    p.out = d.output.data;
    p.os_b = d.output.s0;
    p.os_y = static_cast<i32>(d.output.s1);
    p.os_x = static_cast<i32>(d.output.s2);
    p.wout = static_cast<W*>(d.output_weight.data);
    p.wos_b = d.output_weight.s0;
    p.wos_y = static_cast<i32>(d.output_weight.s1);
    p.wos_x = static_cast<i32>(d.output_weight.s2);
    p.out_h = static_cast<i32>(d.out_shape[2]);
    p.out_wr = static_cast<i32>(d.out_shape[3] / 2 + 1);
    p.out_f[0] = static_cast<C>(d.out_shape[2]);
    p.out_f[1] = static_cast<C>(d.out_shape[3]);
    p.out_centered = d.dst_centered;
    p.n_out = static_cast<i32>(d.out_shape[0]);
    p.tiles_y = static_cast<i32>((d.out_shape[2] + IE_TILE - 1) / IE_TILE);
    p.tiles_x = (p.out_wr + IE_TILE - 1) / IE_TILE;
    p.rot_in = xform_arg<C>(d.rot_in);
    p.scale_in = xform_arg<C>(d.scale_in);
    p.rot_out = xform_arg<C>(d.rot_out);
    p.scale_out = xform_arg<C>(d.scale_out);
    // noa's constructor (:791-845), in the coordinate precision.
    if (d.ews_radius[0] != 0 || d.ews_radius[1] != 0) {
        p.ews[0] = 1 / (2 * static_cast<C>(d.ews_radius[0]));
        p.ews[1] = 1 / (2 * static_cast<C>(d.ews_radius[1]));
    }
    C cutoff = static_cast<C>(d.cutoff);
    cutoff = cutoff > C(0) ? cutoff : C(0);
    p.cutoff_sqd = cutoff * cutoff;
    p.volume_z = static_cast<C>(d.out_shape[2] < d.out_shape[3] ? d.out_shape[2] : d.out_shape[3]);
    const C min_f = 1 / p.volume_z;
    auto clamp_min = [&](double v) { const C c = static_cast<C>(v); return c > min_f ? c : min_f; };
    p.in_sinc = clamp_min(d.in_sinc);
    p.in_blackman = clamp_min(d.in_blackman);
    p.w_sinc = clamp_min(d.w_sinc);
    p.w_blackman = clamp_min(d.w_blackman);
    p.w_size = static_cast<i32>(d.w_size);
    p.add = d.add;
    p.correct = d.correct;
    p.cull = !(d.ews_radius[0] != 0 || d.ews_radius[1] != 0);

    const i64 tiles = static_cast<i64>(p.tiles_y) * p.tiles_x * p.n_out;
    const i64 blocks = (tiles + IE_WAVES - 1) / IE_WAVES;
    if (blocks > 0x7fffffff) { err = "insert_and_extract: too many output slices"; return NOA_HIP_ERROR_UNSUPPORTED; }
    const dim3 grid(static_cast<u32>(blocks)), block(IE_TILE * IE_TILE * IE_WAVES);
    p.interp_mode = d.interp;
    auto launch = [&](auto interp_tag) {
        constexpr int I = decltype(interp_tag)::value;
        if (d.general) hipLaunchKernelGGL((insert_extract_kernel<C, In, Out, W, I, true>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((insert_extract_kernel<C, In, Out, W, I, false>), grid, block, 0, stream, p);
    };
    if (d.interp == NOA_HIP_INTERP_NEAREST) launch(std::integral_constant<int, NOA_HIP_INTERP_NEAREST>{});
    else if (d.interp == NOA_HIP_INTERP_LINEAR) launch(std::integral_constant<int, NOA_HIP_INTERP_LINEAR>{});
    else launch(std::integral_constant<int, INTERP_WINDOW>{});
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class C>
int dispatch(const InsertExtractDesc& d, hipStream_t stream, std::string& err) {
    const int in = d.input.dtype, out = d.output.dtype;
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return run<C, float2, float2, float>(d, stream, err);
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return run<C, float2, float, float>(d, stream, err);
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return run<C, float, float, float>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return run<C, double2, double2, double>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return run<C, double2, double, double>(d, stream, err);
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return run<C, double, double, double>(d, stream, err);
    err = "insert_and_extract: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_insert_extract(const InsertExtractDesc& d, hipStream_t stream, std::string& err) {
    return d.coord_f64 ? dispatch<double>(d, stream, err) : dispatch<float>(d, stream, err);
}

} // namespace noa_hip
