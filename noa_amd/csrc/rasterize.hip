// Fourier insertion of central slices by trilinear rasterisation (scatter).
//
// Replaces FourierInsertRasterize (src/noa/xform/FourierProject.hpp:161-390) launched by
// launch_rasterize_central_slices_3d (:1260-1317).
//
// Design (gfx950): one lane per slice pixel, 256-lane workgroups over consecutive pixels of one slice (the
// slice reads are coalesced and the rotation is a scalar load). Each in-cutoff pixel spreads its value over
// the 2x2x2 trilinear cube (plus the conjugate x=0 duplicates) with device-wide hardware float atomics
// (global_atomic_add_f32) on the coarse-grained volume; the volume is small enough to stay in L2/MALL, where
// the atomics execute. Summation order across pixels is not fixed: the result is deterministic only up to
// float reassociation, as with noa's CUDA and OpenMP paths.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>

namespace noa_hip {

template<class C, class In, class Out, class W>
struct RasterParams {
    const In* in; i64 is_b; i32 is_y, is_x; bool in_is_value; In in_value;
    const W* win; i64 wis_b; i32 wis_y, wis_x; bool win_present, win_is_value; W win_value;
    Out* vol; i32 vs0, vs1, vs2;
    W* wvol; i32 wvs0, wvs1, wvs2; bool has_wvol;
    i32 grid[3];
    C target[3];
    C slice_f[2];
    i32 slice_h, slice_wr, pixels, tiles;
    C cutoff_sqd;
    C ews[2];
    XformArg<C> rot, scale;
    bool general, slices_centered, vol_centered;
};

template<class C, class In, class Out, class W>
__global__ __launch_bounds__(256) void rasterize_kernel(const RasterParams<C, In, Out, W> p) {
    using RIn = typename vt<In>::real;
    using RO = typename vt<Out>::real;
    const u32 lb = xcd_swizzle(blockIdx.x, gridDim.x);
    const i32 b = static_cast<i32>(lb / static_cast<u32>(p.tiles));
    const i32 tile = static_cast<i32>(lb) - b * p.tiles;
    const i32 pix = tile * 256 + static_cast<i32>(threadIdx.x);
    if (pix >= p.pixels) return;
    const i32 y = pix / p.slice_wr;
    const i32 u = pix - y * p.slice_wr;

    const Rot<C> rot = load_rotation(p.rot, b);
    Scale<C> sc;
    sc.identity = true;
    if (p.general) sc = load_scaling(p.scale, b);

    // operator() (:238-267)
    vec2<C> f2;
    f2.y = static_cast<C>(index2frequency(p.slices_centered, y, p.slice_h)) / p.slice_f[0];
    f2.x = static_cast<C>(u) / p.slice_f[1];
    vec3<C> f = p.general ? slice2grid<true>(f2, sc, rot, p.ews[0], p.ews[1])
                          : slice2grid<false>(f2, sc, rot, p.ews[0], p.ews[1]);
    if (dot3(f) > p.cutoff_sqd) return;
    RIn conjugate = 1;
    if (f.x < 0) {
        f.z = -f.z; f.y = -f.y; f.x = -f.x;
        if constexpr (vt<In>::cx) conjugate = -1;
    }
    const C fz = f.z * p.target[0], fy = f.y * p.target[1], fx = f.x * p.target[2];

    // get_input_value_ (:270-277): complex->complex is multiplied by `conjugate`, i.e. NEGATED when -1.
    const In raw = p.in_is_value ? p.in_value : p.in[static_cast<i64>(b) * p.is_b + y * p.is_y + u * p.is_x];
    Out value;
    if constexpr (vt<In>::cx && vt<Out>::cx) value = vcast<Out>(vmul(raw, conjugate));
    else value = vcast<Out>(raw);
    // get_input_weight_ (:279-286)
    W weight{};
    if (p.has_wvol) {
        if (!p.win_present) weight = W(1);
        else weight = p.win_is_value ? p.win_value : p.win[static_cast<i64>(b) * p.wis_b + y * p.wis_y + u * p.wis_x];
    }

    // rasterize_on_3d_grid_ (:309-373)
    const C flz = floor(fz), fly = floor(fy), flx = floor(fx);
    const i32 b0z = static_cast<i32>(flz), b0y = static_cast<i32>(fly), b0x = static_cast<i32>(flx);
    C fr[2][3];
    fr[1][0] = fz - static_cast<C>(b0z);
    fr[1][1] = fy - static_cast<C>(b0y);
    fr[1][2] = fx - static_cast<C>(b0x);
    fr[0][0] = 1 - fr[1][0];
    fr[0][1] = 1 - fr[1][1];
    fr[0][2] = 1 - fr[1][2];
    const i32 D = p.grid[0], H = p.grid[1], width = p.grid[2] / 2 + 1;
#pragma unroll
    for (int w = 0; w < 2; ++w) {
        const i32 iw = frequency2index(p.vol_centered, b0z + w, D);
        if (iw < 0 || iw >= D) continue;
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const i32 iv = frequency2index(p.vol_centered, b0y + v, H);
            if (iv < 0 || iv >= H) continue;
#pragma unroll
            for (int uu = 0; uu < 2; ++uu) {
                const i32 iu = b0x + uu;
                // noa bounds iu by the logical W (:332); taps with iu >= W/2+1 only occur with a zero weight
                // for cutoffs <= 0.5 and would write outside the stored half, so they are skipped.
                if (iu < 0 || iu >= width) continue;
                const C k = fr[w][0] * fr[v][1] * fr[uu][2];
                atomic_add(p.vol + (iw * p.vs0 + iv * p.vs1 + iu * p.vs2), vmul(value, static_cast<RO>(k)));
                if (p.has_wvol)
                    atomic_add(p.wvol + (iw * p.wvs0 + iv * p.wvs1 + iu * p.wvs2), weight * static_cast<W>(k));
            }
        }
    }
    // The x=0 plane is redundant: add the conjugate at (-w, -v, 0) with the same fraction (:347-372).
    if (b0x == 0) {
        if constexpr (vt<Out>::cx) value.y = -value.y;
#pragma unroll
        for (int w = 0; w < 2; ++w) {
            const i32 iw = frequency2index(p.vol_centered, -(b0z + w), D);
            if (iw < 0 || iw >= D) continue;
#pragma unroll
            for (int v = 0; v < 2; ++v) {
                const i32 iv = frequency2index(p.vol_centered, -(b0y + v), H);
                if (iv < 0 || iv >= H) continue;
                const C k = fr[w][0] * fr[v][1] * fr[0][2];
                atomic_add(p.vol + (iw * p.vs0 + iv * p.vs1), vmul(value, static_cast<RO>(k)));
                if (p.has_wvol)
                    atomic_add(p.wvol + (iw * p.wvs0 + iv * p.wvs1), weight * static_cast<W>(k));
            }
        }
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

template<class T> T value_of(const Arr& a) {
    if constexpr (vt<T>::cx) {
        T v; v.x = static_cast<typename vt<T>::real>(a.value[0]); v.y = static_cast<typename vt<T>::real>(a.value[1]);
        return v;
    } else {
        return static_cast<T>(a.value[0]);
    }
}

template<class C, class In, class Out, class W>
int run(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    RasterParams<C, In, Out, W> p{};
    p.in = static_cast<const In*>(d.input.data); p.is_b = d.input.s0; p.is_y = (i32)d.input.s1; p.is_x = (i32)d.input.s2;
    p.in_is_value = d.input.is_value; p.in_value = value_of<In>(d.input);
    p.win_present = d.input_weight.present();
    p.win = static_cast<const W*>(d.input_weight.data);
    p.wis_b = d.input_weight.s0; p.wis_y = (i32)d.input_weight.s1; p.wis_x = (i32)d.input_weight.s2;
    p.win_is_value = d.input_weight.is_value; p.win_value = static_cast<W>(d.input_weight.value[0]);
    p.vol = static_cast<Out*>(d.output.data); p.vs0 = (i32)d.output.s0; p.vs1 = (i32)d.output.s1; p.vs2 = (i32)d.output.s2;
    p.has_wvol = d.output_weight.data != nullptr;
    p.wvol = static_cast<W*>(d.output_weight.data);
    p.wvs0 = (i32)d.output_weight.s0; p.wvs1 = (i32)d.output_weight.s1; p.wvs2 = (i32)d.output_weight.s2;
    for (int i = 0; i < 3; ++i) {
        p.grid[i] = static_cast<i32>(d.out_shape[i + 1]);
        p.target[i] = static_cast<C>(d.target[i]);
    }
    p.slice_f[0] = static_cast<C>(d.in_shape[2]);
    p.slice_f[1] = static_cast<C>(d.in_shape[3]);
    p.slice_h = static_cast<i32>(d.in_shape[2]);
    p.slice_wr = static_cast<i32>(d.in_shape[3] / 2 + 1);
    p.pixels = p.slice_h * p.slice_wr;
    p.tiles = (p.pixels + 255) / 256;
    C cutoff = std::max(static_cast<C>(d.cutoff), C(0));
    p.cutoff_sqd = cutoff * cutoff;
    p.general = d.general;
    if (d.general && (d.ews_radius[0] != 0 || d.ews_radius[1] != 0)) {
        p.ews[0] = 1 / (2 * static_cast<C>(d.ews_radius[0]));
        p.ews[1] = 1 / (2 * static_cast<C>(d.ews_radius[1]));
    }
    p.rot = to_arg<C>(d.rot);
    p.scale = to_arg<C>(d.scale);
    p.slices_centered = d.src_centered;
    p.vol_centered = d.dst_centered;

    const i64 blocks = d.in_shape[0] * p.tiles;
    if (blocks == 0) return NOA_HIP_OK;
    if (blocks > 0x7fffffffLL) { err = "too many slices for one launch"; return NOA_HIP_ERROR_UNSUPPORTED; }
    hipLaunchKernelGGL((rasterize_kernel<C, In, Out, W>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class C>
int dispatch(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    const int in = d.input.dtype, out = d.output.dtype;
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return run<C, float2, float2, float>(d, stream, err);
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return run<C, float2, float, float>(d, stream, err);
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return run<C, float, float, float>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return run<C, double2, double2, double>(d, stream, err);
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return run<C, double2, double, double>(d, stream, err);
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return run<C, double, double, double>(d, stream, err);
    err = "rasterize: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_rasterize(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    return d.coord_f64 ? dispatch<double>(d, stream, err) : dispatch<float>(d, stream, err);
}

} // namespace noa_hip
