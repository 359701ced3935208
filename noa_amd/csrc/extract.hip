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
#include "tiles.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

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
    i32 patch_wgs, patch_pc;            // extract_c32_kernel<PATCH>: workgroups of 8x8 patches per slice, patches per row
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

// PATCH: each wave takes an 8 x 8 patch of pixels (rows y0..y0+7, columns u0..u0+7) instead of 64 consecutive
// pixels of a row: the patch samples a compact piece of the plane, so each tap load touches fewer cache lines of the
// volume. Workgroups [0, patch_wgs) of a slice hold four patches each (columns 0..wr-2); the next ones take the last
// column (u = wr - 1), 256 rows per workgroup.
template<int SHIFT, bool BRICK, class S, bool PATCH = false>
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

    constexpr i32 MAXP = PATCH ? 1 : 4;
    Taps3<float2, float> tv[MAXP];
    bool inside[MAXP];
    i32 ys[MAXP], us[MAXP];
    const i32 first = tile * p.ppt * 256 + static_cast<i32>(threadIdx.x);
#pragma unroll
    for (i32 k = 0; k < MAXP; ++k) {
        bool live;
        i32 y, u;
        if constexpr (PATCH) {
            const i32 wave = static_cast<i32>(threadIdx.x) >> 6, lane = static_cast<i32>(threadIdx.x) & 63;
            if (tile < p.patch_wgs) {
                const i32 q = tile * 4 + wave, pr = q / p.patch_pc, pc = q - pr * p.patch_pc;
                y = pr * 8 + (lane >> 3);
                u = pc * 8 + (lane & 7);
                live = true;
            } else {
                y = (tile - p.patch_wgs) * 256 + static_cast<i32>(threadIdx.x);
                u = p.slice_wr - 1;
                live = y < p.slice_h;
                y = live ? y : 0;
            }
        } else {
            const i32 pix = first + k * 256;
            live = k < p.ppt && pix < p.pixels;
            // pix / wr from a float reciprocal: (pix + 0.5) / wr is at least 0.5 / wr away from an integer, far
            // above the rounding error for pix < 2^21.
            const i32 yq = static_cast<i32>((static_cast<float>(pix) + 0.5f) * p.inv_wr);
            y = live ? yq : 0;
            u = live ? pix - yq * p.slice_wr : 0;
        }
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

// ---------------------------------------------------------------------------------------------------------
// Tiled extraction, for many slices from one volume (cubic volume and slices, Mat33<f32> rotations, no target
// shape, cutoff <= 0.5; the lean path's types and options otherwise). The sample positions (floor of the flipped
// coordinates, in samples) are cut into tiles of 16^3; with cutoff <= 0.5 every pixel inside the cutoff has its
// floor within [lo - 1, hi + 1] along z and y and [0, n/2] along x, which the tiles cover. A work item is (tile,
// range of slices), one workgroup of four waves:
//   - The tile's 17^3 values (taps reach floor + 1; out-of-bounds positions hold 0, which is what noa's bounds
//     checks add for them) are staged in LDS once.
//   - Each wave takes 64 of the item's slices at a time; lane l tests whether slice l's plane crosses the tile, and
//     the crossing slices are processed one at a time.
//   - For a crossing slice, the pixels whose sample falls in the tile (directly, or mirrored by the x < 0 flip) are
//     enumerated row by row: lanes 0-31 take rows of the direct set and lanes 32-63 rows of the mirrored one, each
//     computes its row's pixel range from the tile's box (with a margin far above the rounding), a prefix sum packs
//     the candidates into the wave's LDS list, and every candidate runs noa's expressions: the pixel is taken only
//     where its own floor lies in this tile, so each pixel inside the cutoff is computed by exactly one item with
//     the same arithmetic as the gather kernels (bit-identical), its taps read from LDS.
//   - The pixels outside the cutoff (value 0, times the phase shift) are written by extract_outside_kernel.
// The volume is read about once (plus the tiles' halo), instead of once per slice footprint.
// ---------------------------------------------------------------------------------------------------------
struct TileParams {
    const void* vol; i32 vs0, vs1;      // stored volume, unit x stride
    float2* out; i64 os_b; i32 os_y;    // output slices, unit x stride
    XformArg<float> rot;
    const int4* items;                  // {tile, first slice, end slice, 0}
    i32 n;                              // logical size of the volume's and the slices' axes
    i32 ntz, nty, ntx;                  // tiles along z, y, x
    i32 base;                           // floor position of the first tile along z and y
    i32 n_slices, pixels, tiles_px;     // pixels per slice; workgroups of 256 pixels per slice (zero pass)
    float inv_wr;
    float cutoff_sqd, margin;
    bool vol_centered, slice_centered;
    const float* shifts;                // per-slice (y, x) shifts, or null (broadcast shift_f)
    float shift_f[2], norm_f[2], cutoff_sqd_f;
};

template<int SHIFT>
__device__ __forceinline__ float2 tile_shift(const TileParams& p, i32 b, i32 fy, i32 u, float2 value) {
    if constexpr (SHIFT == 0) {
        return value;
    } else {
        float sy = 0.f, sx = 0.f;
        if constexpr (SHIFT == 1) {
            sy = p.shifts ? p.shifts[2 * b] : p.shift_f[0];
            sx = p.shifts ? p.shifts[2 * b + 1] : p.shift_f[1];
        }
        return cmul(value, phase_shift_at<float2, float>(SHIFT == 1 ? 0 : 2, sy, sx, fy, u, p.norm_f[0], p.norm_f[1],
                                                          p.cutoff_sqd_f));
    }
}


template<int SHIFT, class S>
__global__ __launch_bounds__(256) void extract_tile_kernel(const TileParams p) {
    __shared__ float2 tile[XTT * XTT * XTT];
    __shared__ u32 list[XT_WAVES][XT_LIST];

    const int4 item = p.items[blockIdx.x];
    const i32 first = item.y, end = min(item.z, p.n_slices);
    const i32 tx = item.x % p.ntx, tzy = item.x / p.ntx;
    const i32 ty = tzy % p.nty, tz = tzy / p.nty;
    const i32 Z0 = p.base + tz * XT, Y0 = p.base + ty * XT, X0 = tx * XT;
    const i32 n = p.n, lo = -n / 2, hi = (n - 1) / 2, hix = n / 2;

    // Stage the tile: positions [Z0, Z0 + 16] x [Y0, Y0 + 16] x [X0, X0 + 16], zero outside the stored frequencies.
    const S* __restrict__ vol = static_cast<const S*>(p.vol);
    for (i32 i = static_cast<i32>(threadIdx.x); i < XTT * XTT * XTT; i += 256) {
        const i32 dz = i / (XTT * XTT), r = i - dz * (XTT * XTT);
        const i32 dy = r / XTT, dx = r - dy * XTT;
        const i32 z = Z0 + dz, y = Y0 + dy, x = X0 + dx;
        float2 v = make_float2(0.f, 0.f);
        if (z >= lo && z <= hi && y >= lo && y <= hi && x <= hix)
            v = load_as<float2>(vol + (frequency2index(p.vol_centered, z, n) * p.vs0 +
                                       frequency2index(p.vol_centered, y, n) * p.vs1 + x));
        tile[i] = v;
    }
    __syncthreads();

    const i32 wave = static_cast<i32>(threadIdx.x) >> 6, lane = static_cast<i32>(threadIdx.x) & 63;
    u32* wl = list[wave];
    const float slice_f = static_cast<float>(n), target = static_cast<float>(n);
    const float M = p.margin;
    // Tile centre and half-extent in samples (floor positions [Z0, Z0 + 16) are samples in [Z0, Z0 + 16)).
    const float cz = static_cast<float>(Z0) + 0.5f * XT, cy = static_cast<float>(Y0) + 0.5f * XT;
    const float cx = static_cast<float>(X0) + 0.5f * XT;
    Scale<float> sc;
    sc.identity = true;

    for (i32 c0 = first + wave * 64; c0 < end; c0 += 64 * XT_WAVES) {
        // Does the plane of slice c0 + lane (spanned by the rotation's 2nd and 3rd columns) cross the tile box? The
        // plane is symmetric through the origin, so this also decides the mirrored box.
        const bool cross = c0 + lane < end && tile_crossed(load_rotation(p.rot, c0 + lane), cz, cy, cx, M);
        unsigned long long mask = __ballot(cross);
        while (mask) {
            const i32 k = __ffsll(static_cast<long long>(mask)) - 1;
            mask &= mask - 1;
            const i32 b = c0 + k;
            const Rot<float> r = load_rotation(p.rot, b);
            float2* __restrict__ out = p.out + static_cast<i64>(b) * p.os_b;

            // Evaluates one candidate pixel: noa's expressions (as extract_c32_kernel), kept only when its floor is
            // in this tile and its flip matches the set it was enumerated from.
            // A candidate in two phases, so that a lane can have the LDS reads of two candidates in flight: fetch runs
            // noa's expressions (as extract_c32_kernel) and reads the taps, and keeps the candidate only when its
            // floor is in this tile and its flip matches the set it was enumerated from; store combines and writes.
            struct Cand {
                Taps3<float2, float> t;
                i32 fy, u;
                bool ok;
            };
            auto fetch = [&](i32 fy, i32 u, bool mirrored, Cand& c) {
                c.ok = false;
                c.fy = fy;
                c.u = u;
                vec2<float> f2;
                f2.y = static_cast<float>(fy) / slice_f;
                f2.x = static_cast<float>(u) / slice_f;
                const vec3<float> f = slice2grid<false>(f2, sc, r, 0.f, 0.f);
                if (!(dot3(f) <= p.cutoff_sqd)) return;
                float fz = f.z * target, fyv = f.y * target, fx = f.x * target;
                c.t.conj = 1;
                const bool flip = fx < 0;
                if (flip != mirrored) return;
                if (flip) { fz *= -1; fyv *= -1; fx *= -1; c.t.conj = -1; }
                const float flz = floorf(fz), fly = floorf(fyv), flx = floorf(fx);
                const i32 lz = static_cast<i32>(flz) - Z0, ly = static_cast<i32>(fly) - Y0, lx = static_cast<i32>(flx) - X0;
                if (static_cast<u32>(lz) >= static_cast<u32>(XT) || static_cast<u32>(ly) >= static_cast<u32>(XT) ||
                    static_cast<u32>(lx) >= static_cast<u32>(XT))
                    return;
                c.t.tz = fz - flz;
                c.t.ty = fyv - fly;
                c.t.tx = fx - flx;
                c.t.ok = 0xffu;
                const i32 o = (lz * XTT + ly) * XTT + lx;
#pragma unroll
                for (int z = 0; z < 2; ++z)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
#pragma unroll
                        for (int x = 0; x < 2; ++x) c.t.v[z][y][x] = tile[o + (z * XTT + y) * XTT + x];
                c.ok = true;
            };
            auto store = [&](const Cand& c) {
                if (!c.ok) return;
                float2 value = combine3<NOA_HIP_INTERP_LINEAR>(c.t);
                value = tile_shift<SHIFT>(p, b, c.fy, c.u, value);
                out[frequency2index(p.slice_centered, c.fy, n) * p.os_y + c.u] = value;
            };

            tile_candidates<Cand>(r, Z0, Y0, X0, n, M, p.pixels, p.inv_wr, p.slice_centered, lane, wl, fetch, store);
        }
    }
}

// The pixels outside the cutoff of the tiled extraction: 0, times the phase shift as the gather kernels compute it.
// Four pixels per lane (consecutive groups of 256). The transform stretches a slice frequency f by at most
// sqrt(lambda), lambda the largest eigenvalue of G^T G (G: its 2nd and 3rd columns); a pixel with lambda |f|^2 inside
// the cutoff by a margin far above the rounding is inside after the transform too and is skipped.
template<int SHIFT>
__global__ __launch_bounds__(256) void extract_outside_kernel(const TileParams p) {
    const u32 bid = blockIdx.x;
    const i32 b = static_cast<i32>(bid / static_cast<u32>(p.tiles_px));
    const i32 first = (static_cast<i32>(bid) - b * p.tiles_px) * 1024 + static_cast<i32>(threadIdx.x);
    const i32 n = p.n, wr = n / 2 + 1;
    const float inv_n = 1.f / static_cast<float>(n);
    const Rot<float> r = load_rotation(p.rot, b);
    const float aa = (r.m[1] * r.m[1] + r.m[4] * r.m[4]) + r.m[7] * r.m[7];
    const float cc = (r.m[2] * r.m[2] + r.m[5] * r.m[5]) + r.m[8] * r.m[8];
    const float ac = (r.m[1] * r.m[2] + r.m[4] * r.m[5]) + r.m[7] * r.m[8];
    const float hd = 0.5f * (aa - cc);
    const float lambda = 0.5f * (aa + cc) + sqrtf(hd * hd + ac * ac);
    const float inner = p.cutoff_sqd * (1.f - 1e-4f) / fmaxf(lambda * (1.f + 1e-4f), 1e-30f);
    bool any = false;
    i32 yqs[4], us[4], fys[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const i32 pix = first + k * 256;
        yqs[k] = -1;
        if (pix >= p.pixels) continue;
        const i32 yq = static_cast<i32>((static_cast<float>(pix) + 0.5f) * p.inv_wr);
        const i32 u = pix - yq * wr;
        const i32 fy = index2frequency(p.slice_centered, yq, n);
        const float gy = static_cast<float>(fy) * inv_n, gx = static_cast<float>(u) * inv_n;
        if (gy * gy + gx * gx < inner) continue;
        yqs[k] = yq;
        us[k] = u;
        fys[k] = fy;
        any = true;
    }
    if (!any) return;
    Scale<float> sc;
    sc.identity = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (yqs[k] < 0) continue;
        vec2<float> f2;
        f2.y = static_cast<float>(fys[k]) / static_cast<float>(n);
        f2.x = static_cast<float>(us[k]) / static_cast<float>(n);
        const vec3<float> f = slice2grid<false>(f2, sc, r, 0.f, 0.f);
        if (dot3(f) <= p.cutoff_sqd) continue;   // computed by the tile kernel
        p.out[static_cast<i64>(b) * p.os_b + yqs[k] * p.os_y + us[k]] =
            tile_shift<SHIFT>(p, b, fys[k], us[k], vzero<float2>());
    }
}

// ---------------------------------------------------------------------------------------------------------
// Batched extraction: several extraction calls (each its own volume, box size, layout, slice count and cutoff) in
// one persistent launch. The device image of a batch holds a header and one ExtractBatchJob per call; a call's
// slices are cut into units of ppt * 256 consecutive pixels of one slice (ppt from the header), numbered across
// the calls. The grid is fixed (any size works, so a launch can be captured once in a graph and replayed for
// other batches): workgroup g takes the contiguous units [U g / G, U (g + 1) / G), the workgroups of one XCD
// neighbouring ranges (mostly one volume per XCD L2). Each pixel runs the same expressions as extract_c32_kernel,
// so the results are those of the calls made one by one.
// ---------------------------------------------------------------------------------------------------------
struct ExtractBatchJob {
    const float2* vol; float2* out;
    XformArg<float> rot;
    i64 first_unit;                     // index of the call's first unit in the batch
    i64 os_b; i32 os_y;                 // output strides (unit x stride)
    i32 vs0, vs1;                       // volume strides (unit x stride)
    i32 grid[3];                        // logical D, H, W of the volume
    i32 slice_h, slice_wr;
    i32 pixels, tiles;                  // pixel slots of a slice (8 x 8 patches, then the last column) and units
    i32 patch_cols, patch_pixels;       // patches per row of patches (columns 0 .. wr - 2), and their pixel slots
    float target[3], slice_f[2];
    float cutoff_sqd;
    i32 vol_centered, slice_centered;
    i32 blocks;                         // the volume is a block copy (noa_hip_volume_blocks): vs0, vs1 are its strides
};
struct ExtractBatchHeader {
    i64 n_units;
    i32 n_jobs, ppt;
    i32 pad[12];                        // jobs start at byte 64
};
static_assert(sizeof(ExtractBatchHeader) == 64, "batch header");

__global__ __launch_bounds__(256) void extract_batch_kernel(const unsigned char* __restrict__ image) {
    const ExtractBatchHeader* h = reinterpret_cast<const ExtractBatchHeader*>(image);
    const ExtractBatchJob* __restrict__ jobs = reinterpret_cast<const ExtractBatchJob*>(image + sizeof(ExtractBatchHeader));
    const i64 U = h->n_units;
    const i32 J = h->n_jobs, ppt = h->ppt;
    const u32 g = xcd_swizzle(blockIdx.x, gridDim.x);
    const i64 u_begin = U * g / gridDim.x, u_end = U * (g + 1) / gridDim.x;
    if (u_begin >= u_end) return;
    i32 lo = 0, hi = J - 1;                              // last job whose first unit is <= u_begin
    while (lo < hi) {
        const i32 mid = (lo + hi + 1) / 2;
        if (jobs[mid].first_unit <= u_begin) lo = mid;
        else hi = mid - 1;
    }
    i32 j = lo;
    constexpr i32 MAXP = 4;
    for (i64 u = u_begin; u < u_end; ++u) {
        while (j + 1 < J && jobs[j + 1].first_unit <= u) ++j;
        const ExtractBatchJob& q = jobs[j];
        const i32 local = static_cast<i32>(u - q.first_unit);
        const i32 b = local / q.tiles, tile = local - b * q.tiles;
        const Rot<float> rot = load_rotation(q.rot, b);
        Scale<float> sc;
        sc.identity = true;
        const Bounds3 bd = make_bounds3(q.grid[0], q.grid[1], q.grid[2]);
        const bool vc = q.vol_centered != 0, scn = q.slice_centered != 0;
        float2* __restrict__ out = q.out + static_cast<i64>(b) * q.os_b;
        Taps3<float2, float> tv[MAXP];
        bool inside[MAXP];
        i32 ys[MAXP], us[MAXP];
        const i32 first = tile * ppt * 256 + static_cast<i32>(threadIdx.x);
#pragma unroll
        for (i32 k = 0; k < MAXP; ++k) {
            // Pixel slots: 8 x 8 patches of columns 0 .. wr - 2 (a wave's 64 slots are one patch, a compact piece of
            // the plane: fewer cache lines per tap load than 64 pixels of a row), then the last column.
            const i32 pix = first + k * 256;
            const bool in_range = k < ppt && pix < q.pixels;
            i32 y, uu;
            bool live;
            if (pix < q.patch_pixels) {                  // uniform per wave: patch_pixels is a multiple of 64
                const i32 patch = pix >> 6, l = pix & 63;
                const i32 pr = patch / q.patch_cols, pc = patch - pr * q.patch_cols;
                y = pr * 8 + (l >> 3);
                uu = pc * 8 + (l & 7);
                live = in_range && y < q.slice_h && uu < q.slice_wr - 1;
            } else {
                y = pix - q.patch_pixels;
                uu = q.slice_wr - 1;
                live = in_range && y < q.slice_h;
            }
            y = live ? y : 0;
            uu = live ? uu : 0;
            ys[k] = live ? y : -1;
            us[k] = uu;
            vec2<float> f2;
            const i32 fy = index2frequency(scn, y, q.slice_h);
            f2.y = static_cast<float>(fy) / q.slice_f[0];
            f2.x = static_cast<float>(uu) / q.slice_f[1];
            const vec3<float> f = slice2grid<false>(f2, sc, rot, 0.f, 0.f);
            inside[k] = live && dot3(f) <= q.cutoff_sqd;
            const float fz = f.z * q.target[0], fyv = f.y * q.target[1], fx = f.x * q.target[2];
            if (q.blocks)
                tv[k] = fetch3<NOA_HIP_INTERP_LINEAR, float2>(GlobalLoad<float2>{q.vol}, BrickLayout{q.vs0, q.vs1}, bd, vc,
                                                              fz, fyv, fx, inside[k]);
            else
                tv[k] = fetch3<NOA_HIP_INTERP_LINEAR, float2>(GlobalLoad<float2>{q.vol}, q.vs0, q.vs1, 1, bd, vc, fz, fyv,
                                                              fx, inside[k]);
        }
#pragma unroll
        for (i32 k = 0; k < MAXP; ++k) {
            if (ys[k] < 0) continue;
            const float2 value = inside[k] ? combine3<NOA_HIP_INTERP_LINEAR>(tv[k]) : vzero<float2>();
            out[ys[k] * q.os_y + us[k]] = value;
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
            // Tiled path (extract_tile_kernel): cubic volume and slices, no target shape, cutoff <= 0.5, Mat33<f32>
            // rotations; chosen by the tuning table (layout 3) or NOA_HIP_EXTRACT_LAYOUT=tiles.
            // Launch tuning: volume layout and tile work-item size; NOA_HIP_EXTRACT_LAYOUT overrides the table.
            i32 tv[4];
            tuning_lookup(NOA_HIP_OP_EXTRACT, d.in_shape[1], (d.src_centered ? 4 : 0) + (d.dst_centered ? 8 : 0),
                          d.interp, d.input.dtype, tv);
            static const char* const LAYOUTS[4] = {nullptr, "linear", "bricks", "tiles"};
            const char* layout_env = std::getenv("NOA_HIP_EXTRACT_LAYOUT");
            if (!layout_env && tv[0] >= 1 && tv[0] <= 3) layout_env = LAYOUTS[tv[0]];
            const i64 nn = p.grid[0];
            const bool cubic = p.grid[1] == nn && p.grid[2] == nn && d.out_shape[2] == nn && d.out_shape[3] == nn &&
                               d.target[0] == nn && d.target[1] == nn && d.target[2] == nn;
            const i64 stored_elems = nn * nn * (nn / 2 + 1);
            // Only on request: at 256^3 with 4096 slices the block-copy gather is faster (1.15 against 1.74 ms).
            bool tiled = false;
            if (layout_env && std::strcmp(layout_env, "tiles") == 0)
                tiled = cubic && nn >= 16 && nn <= 1000 && static_cast<double>(cutoff) <= 0.5 && p.os_x == 1 &&
                        p.rot.kind == NOA_HIP_MAT33_F32 && stored_elems < 0x7fffffffLL;
            const i64 tiles_px = (p.pixels + 1023) / 1024;   // outside kernel: 1024 pixels per workgroup
            if (tiled && n_slices * tiles_px <= 0x7fffffffLL) {
                int device = 0;
                (void)hipGetDevice(&device);
                // Plans are kept per slice count rounded up to m * 2^k, 8 <= m < 16 (items are clamped to the
                // actual count in the kernel).
                i64 step = 1;
                while (n_slices > 16 * step) step *= 2;
                const i32 plan_slices = static_cast<i32>((n_slices + step - 1) / step * step);
                const char* pe = std::getenv("NOA_HIP_EXTRACT_PAIRS");
                const i32 pairs = pe ? std::max(8, std::atoi(pe)) : (tv[1] > 0 ? std::max(8, tv[1]) : 128);
                TilePlan plan;
                const int st = make_tile_plan(device, static_cast<i32>(nn), static_cast<double>(cutoff), plan_slices,
                                              pairs, plan, err);
                if (st != NOA_HIP_OK) return st;
                TileParams tp{};
                tp.vol = p.vol; tp.vs0 = p.vs0; tp.vs1 = p.vs1;
                tp.out = static_cast<float2*>(p.out); tp.os_b = p.os_b; tp.os_y = p.os_y;
                tp.rot = p.rot;
                tp.items = plan.items;
                tp.n = static_cast<i32>(nn);
                tp.ntz = plan.ntz; tp.nty = plan.nty; tp.ntx = plan.ntx; tp.base = plan.base;
                tp.n_slices = static_cast<i32>(n_slices);
                tp.pixels = p.pixels;
                tp.tiles_px = static_cast<i32>(tiles_px);
                tp.inv_wr = p.inv_wr;
                tp.cutoff_sqd = static_cast<float>(p.cutoff_sqd);
                tp.margin = 1e-3f + 1e-5f * static_cast<float>(nn);
                tp.vol_centered = p.vol_centered;
                tp.slice_centered = p.slice_centered;
                tp.shifts = p.shift_mode == 0 && p.shift_batched ? static_cast<const float*>(p.shifts) : nullptr;
                for (int i = 0; i < 2; ++i) { tp.shift_f[i] = p.shift_f[i]; tp.norm_f[i] = p.norm_f[i]; }
                tp.cutoff_sqd_f = p.cutoff_sqd_f;
                const dim3 gi(static_cast<u32>(plan.n_items)), go(static_cast<u32>(n_slices * tiles_px));
                if (p.shift_mode == 0) {
                    if (plan.n_items > 0) hipLaunchKernelGGL((extract_tile_kernel<1, S>), gi, block, 0, stream, tp);
                    hipLaunchKernelGGL((extract_outside_kernel<1>), go, block, 0, stream, tp);
                } else if (p.shift_mode == 2) {
                    if (plan.n_items > 0) hipLaunchKernelGGL((extract_tile_kernel<2, S>), gi, block, 0, stream, tp);
                    hipLaunchKernelGGL((extract_outside_kernel<2>), go, block, 0, stream, tp);
                } else {
                    if (plan.n_items > 0) hipLaunchKernelGGL((extract_tile_kernel<0, S>), gi, block, 0, stream, tp);
                    hipLaunchKernelGGL((extract_outside_kernel<0>), go, block, 0, stream, tp);
                }
                const hipError_t e = hipGetLastError();
                if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
                return NOA_HIP_OK;
            }
            // Many slices from one volume: gather from a block copy of it (BrickLayout) (one read and one write of the
            // volume, then every trilinear gather touches fewer cache lines). Worth it once the slices hold about
            // twice the volume's elements; NOA_HIP_EXTRACT_LAYOUT=linear|bricks forces either.
            const i32 nz = p.grid[0], ny = p.grid[1], wr = p.grid[2] / 2 + 1;
            const i64 stored = static_cast<i64>(nz) * ny * wr;
            i64 bz, by, brick_elems;
            BrickLayout::strides(nz, ny, wr, bz, by, brick_elems);
            const char* env = layout_env;
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
            // Pixel patches (extract_c32_kernel<PATCH>): rows and the columns but the last in blocks of 8, four
            // patches per workgroup.
            const i32 pr = p.slice_h / 8, pc = (p.slice_wr - 1) / 8;
            bool patch = p.slice_h % 8 == 0 && (p.slice_wr - 1) % 8 == 0 && (pr * pc) % 4 == 0 && pr * pc > 0;
            if (const char* pe = std::getenv("NOA_HIP_EXTRACT_PATCH")) patch = patch && pe[0] != '0';
            const i64 patch_tiles = pr * pc / 4 + (p.slice_h + 255) / 256;
            patch = patch && n_slices * patch_tiles <= 0x7fffffffLL;
            if (brick) {
                const BrickLayout L{static_cast<i32>(bz), static_cast<i32>(by)};
                hipLaunchKernelGGL((brick_volume_kernel<S>), dim3(static_cast<u32>((stored + 255) / 256)), dim3(256), 0,
                                   stream, static_cast<const S*>(p.vol), p.vs0, p.vs1, static_cast<S*>(brick),
                                   ny, wr, stored, L);
                ExtractParams<C> q = p;
                q.vol = brick;
                q.vs0 = L.bz;
                q.vs1 = L.by;
                if (patch) {
                    q.ppt = 1;
                    q.patch_wgs = pr * pc / 4;
                    q.patch_pc = pc;
                    q.tiles = static_cast<i32>(patch_tiles);
                    const dim3 pgrid(static_cast<u32>(n_slices * patch_tiles));
                    if (q.shift_mode == 0) hipLaunchKernelGGL((extract_c32_kernel<1, true, S, true>), pgrid, block, 0, stream, q);
                    else if (q.shift_mode == 2) hipLaunchKernelGGL((extract_c32_kernel<2, true, S, true>), pgrid, block, 0, stream, q);
                    else hipLaunchKernelGGL((extract_c32_kernel<0, true, S, true>), pgrid, block, 0, stream, q);
                } else {
                    if (q.shift_mode == 0) hipLaunchKernelGGL((extract_c32_kernel<1, true, S>), grid, block, 0, stream, q);
                    else if (q.shift_mode == 2) hipLaunchKernelGGL((extract_c32_kernel<2, true, S>), grid, block, 0, stream, q);
                    else hipLaunchKernelGGL((extract_c32_kernel<0, true, S>), grid, block, 0, stream, q);
                }
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

i64 extract_batch_bytes(i32 n_jobs) {
    return static_cast<i64>(sizeof(ExtractBatchHeader)) + static_cast<i64>(n_jobs) * static_cast<i64>(sizeof(ExtractBatchJob));
}

int extract_batch_prepare(const ProjectDesc* descs, i32 n_jobs, void* image, i64 capacity, i32 ppt, i64& bytes,
                          std::string& err) {
    bytes = extract_batch_bytes(n_jobs);
    if (ppt < 1 || ppt > 4) { err = "extract batch: pixels per lane must be 1 to 4"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    if (!image || capacity < bytes) {
        err = "extract batch: the image needs " + std::to_string(bytes) + " bytes";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    std::vector<unsigned char> buf(static_cast<size_t>(bytes), 0);
    auto* h = reinterpret_cast<ExtractBatchHeader*>(buf.data());
    auto* jobs = reinterpret_cast<ExtractBatchJob*>(buf.data() + sizeof(ExtractBatchHeader));
    i64 units = 0;
    for (i32 k = 0; k < n_jobs; ++k) {
        const ProjectDesc& d = descs[k];
        const std::string at = "extract batch, call " + std::to_string(k) + ": ";
        if ((d.input.dtype != NOA_HIP_C32 && d.input.dtype != NOA_HIP_C32_BLOCKS) || d.output.dtype != NOA_HIP_C32 ||
            d.coord_f64 || d.general ||
            d.interp != NOA_HIP_INTERP_LINEAR || d.input_weight.data || d.output_weight.data || d.input.s2 != 1 ||
            d.output.s2 != 1 || d.shift.present) {
            err = at + "only complex FP32 volumes and slices with unit x strides, LINEAR interpolation, single-precision "
                       "transforms, no weights, scaling, Ewald sphere or shifts are batched";
            return NOA_HIP_ERROR_UNSUPPORTED;
        }
        const float t0 = static_cast<float>(d.target[0]);
        const float fb = std::max(static_cast<float>(d.fftfreq_blackman), 1 / t0);
        if (blackman_window_size(static_cast<double>(fb), static_cast<double>(t0)) != 1) {
            err = at + "a w-windowed sinc wider than one sample is not batched";
            return NOA_HIP_ERROR_UNSUPPORTED;
        }
        ExtractBatchJob& q = jobs[k];
        q.vol = static_cast<const float2*>(d.input.data);
        q.out = static_cast<float2*>(d.output.data);
        q.rot = to_arg<float>(d.rot);
        q.first_unit = units;
        q.os_b = d.output.s0;
        q.os_y = static_cast<i32>(d.output.s1);
        q.vs0 = static_cast<i32>(d.input.s0);
        q.vs1 = static_cast<i32>(d.input.s1);
        q.blocks = d.input.dtype == NOA_HIP_C32_BLOCKS ? 1 : 0;
        if (q.blocks) {
            i64 bz, by, count;
            BrickLayout::strides(d.in_shape[1], d.in_shape[2], d.in_shape[3] / 2 + 1, bz, by, count);
            if (count >= 0x7fffffffLL) { err = at + "the block copy exceeds 32-bit indexing"; return NOA_HIP_ERROR_UNSUPPORTED; }
            q.vs0 = static_cast<i32>(bz);
            q.vs1 = static_cast<i32>(by);
        }
        for (int i = 0; i < 3; ++i) {
            q.grid[i] = static_cast<i32>(d.in_shape[i + 1]);
            q.target[i] = static_cast<float>(d.target[i]);
        }
        q.slice_h = static_cast<i32>(d.out_shape[2]);
        q.slice_wr = static_cast<i32>(d.out_shape[3] / 2 + 1);
        if (static_cast<i64>(q.slice_h) * q.slice_wr >= (1 << 21)) {
            err = at + "slices above 2^21 pixels are not batched";
            return NOA_HIP_ERROR_UNSUPPORTED;
        }
        q.patch_cols = (q.slice_wr - 1 + 7) / 8;
        q.patch_pixels = (q.slice_h + 7) / 8 * q.patch_cols * 64;
        q.pixels = q.patch_pixels + q.slice_h;
        q.tiles = (q.pixels + 256 * ppt - 1) / (256 * ppt);
        q.slice_f[0] = static_cast<float>(d.out_shape[2]);
        q.slice_f[1] = static_cast<float>(d.out_shape[3]);
        const float cutoff = std::max(static_cast<float>(d.cutoff), 0.f);
        q.cutoff_sqd = cutoff * cutoff;
        q.vol_centered = d.src_centered ? 1 : 0;
        q.slice_centered = d.dst_centered ? 1 : 0;
        units += d.out_shape[0] * q.tiles;
    }
    h->n_units = units;
    h->n_jobs = n_jobs;
    h->ppt = ppt;
    std::memcpy(image, buf.data(), buf.size());
    return NOA_HIP_OK;
}

i64 volume_blocks_bytes(const int64_t rfft_shape[4]) {
    i64 bz, by, count;
    BrickLayout::strides(rfft_shape[1], rfft_shape[2], rfft_shape[3], bz, by, count);
    return count * static_cast<i64>(sizeof(float2));
}

int volume_blocks(const void* volume, i64 s_z, i64 s_y, const int64_t rfft_shape[4], void* blocks, hipStream_t stream,
                  std::string& err) {
    const i64 nz = rfft_shape[1], ny = rfft_shape[2], wr = rfft_shape[3];
    i64 bz, by, count;
    BrickLayout::strides(nz, ny, wr, bz, by, count);
    const i64 stored = nz * ny * wr;
    if (count >= 0x7fffffffLL || s_z * (nz - 1) + s_y * (ny - 1) + wr >= 0x7fffffffLL) {
        err = "volume blocks: the volume exceeds 32-bit indexing";
        return NOA_HIP_ERROR_UNSUPPORTED;
    }
    if (stored == 0) return NOA_HIP_OK;
    const BrickLayout L{static_cast<i32>(bz), static_cast<i32>(by)};
    hipLaunchKernelGGL((brick_volume_kernel<float2>), dim3(static_cast<u32>((stored + 255) / 256)), dim3(256), 0, stream,
                       static_cast<const float2*>(volume), static_cast<i32>(s_z), static_cast<i32>(s_y),
                       static_cast<float2*>(blocks), static_cast<i32>(ny), static_cast<i32>(wr), stored, L);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

int extract_batch_launch(const void* image, i32 grid, hipStream_t stream, std::string& err) {
    if (grid < 1) { err = "extract batch: the grid needs at least one workgroup"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    hipLaunchKernelGGL(extract_batch_kernel, dim3(static_cast<u32>(grid)), dim3(256), 0, stream,
                       static_cast<const unsigned char*>(image));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

} // namespace noa_hip
