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
//
// Many slices into one cubic volume (FP32 values and weights, Mat33<f32> rotations, cutoff <= 0.5) take the tiled
// path instead, rasterize_tile_kernel: the accumulation happens in LDS and reaches the volume once per work item.
#include "common.hpp"
#include "kernels.hpp"
#include "tiles.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>

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

// ---------------------------------------------------------------------------------------------------------
// Tiled rasterisation. The floor positions of the (flipped) pixel coordinates are cut into tiles of 16^3 (as for the
// tiled extraction, tiles.hpp). A work item is (tile, range of slices), one workgroup of four waves; wave w owns the
// slab of floor positions z in [Z0 + 4w, Z0 + 4w + 4) of the tile:
//   - LDS holds each wave's own accumulators over its slab and the layer above it (the trilinear taps reach
//     floor + 1), 5 x 17 x 17 positions of value and weight, zeroed at the start. No other wave touches them.
//   - The wave visits every slice of the item whose plane crosses its slab and enumerates the slab's candidate
//     pixels (tile_candidates). A candidate runs noa's expressions of rasterize_kernel and is kept only where its
//     floor lies in the slab, so every pixel inside the cutoff is rasterised by exactly one wave of one item.
//   - The eight taps are added with plain LDS read-modify-write, tap by tap in program order (LDS executes one
//     wave's instructions in order; compiler barriers keep the order). Lanes whose candidates share a floor
//     position would update the same words in one instruction: each lane writes its id into a tag byte of its
//     floor position and reads it back, the lane that reads its own id does the read-modify-write and the others
//     add with the LDS float atomic (ds_add_f32) after it. Two pixels share a floor position only when they are
//     neighbours in the slice, so these lanes are few; the float atomic costs about two clocks per lane.
//   - The x = 0 duplicates (conjugate at (-w, -v, 0)) go straight to the volume with device-wide atomics, as in
//     rasterize_kernel: they land anywhere in the x = 0 plane and are a thin set.
//   - After the item's slices, each wave adds its 5 x 17 x 17 sums to the volume with device-wide atomics at noa's
//     indices: frequency2index of the position, dropped outside [0, D) x [0, H) x [0, W/2+1) (the bounds checks
//     rasterize_kernel applies per tap, so the wrapped taps of the non-centred layout land where they land there).
//     Neighbouring slabs and tiles overlap by one layer; both add theirs.
// Summation order is not fixed (the device atomics), as with rasterize_kernel.
// ---------------------------------------------------------------------------------------------------------
constexpr i32 RS_Z = 4;                             // floor layers per wave
constexpr i32 RS_L = RS_Z + 1;                      // LDS layers per wave
constexpr i32 RS_CELLS = RS_L * XTT * XTT;          // accumulator positions per wave
constexpr i32 RS_TAGS = RS_Z * XT * XT;             // floor positions per wave
constexpr i32 RS_LIST = 384;                        // candidate pixels per wave and round
static_assert(RS_Z * XT_WAVES == XT, "the waves' slabs cover the tile");

template<class In, class Out>
struct RasterTileParams {
    const In* in; i64 is_b; i32 is_y, is_x; bool in_is_value; In in_value;
    const float* win; i64 wis_b; i32 wis_y, wis_x; bool win_present, win_is_value; float win_value;
    Out* vol; i32 vs0, vs1, vs2;
    float* wvol; i32 wvs0, wvs1, wvs2; bool has_wvol;
    XformArg<float> rot, scale;         // scale: Mat22 inverse scaling, or none
    const int4* items;                  // {tile, first slice, end slice, 0}
    i32 n, nty, ntx, base, n_slices, pixels;
    float inv_wr, cutoff_sqd, margin;
    bool vol_centered, slice_centered, scaled;
};

__device__ __forceinline__ void lds_order() { asm volatile("" ::: "memory"); }

// The plane geometry of the tiled path (crossing test, candidate enumeration) takes the rotation as a matrix; the
// per-pixel expressions keep the rotation as given (noa's quaternion rotation for quaternions).
__device__ __forceinline__ Rot<float> as_matrix(const Rot<float>& r) {
    Rot<float> g;
    rot_matrix(r, g.m);
    g.quat = false;
    return g;
}

// With an inverse scaling S (applied to the 2D frequency before the rotation), pixel (y, x) lands at
// y (s0 a + s2 c) + x (s1 a + s3 c), with a and c the rotation's 2nd and 3rd columns: the same plane, other columns.
__device__ __forceinline__ Rot<float> plane_matrix(const Rot<float>& r, const Scale<float>& s) {
    Rot<float> g = as_matrix(r);
    if (s.identity) return g;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float a = g.m[3 * i + 1], c = g.m[3 * i + 2];
        g.m[3 * i + 1] = s.m[0] * a + s.m[2] * c;
        g.m[3 * i + 2] = s.m[1] * a + s.m[3] * c;
    }
    return g;
}

__device__ __forceinline__ Scale<float> slice_scaling(const XformArg<float>& a, bool scaled, i64 b) {
    Scale<float> s;
    s.identity = true;
    if (scaled) s = load_scaling(a, b);
    return s;
}

template<class In, class Out>
__global__ __launch_bounds__(256) void rasterize_tile_kernel(const RasterTileParams<In, Out> p) {
    using RIn = typename vt<In>::real;
    constexpr bool CX = vt<Out>::cx;
    __shared__ Out vacc_all[XT_WAVES][RS_CELLS];
    __shared__ float wacc_all[XT_WAVES][RS_CELLS];
    __shared__ unsigned char tags_all[XT_WAVES][RS_TAGS];
    __shared__ u32 list[XT_WAVES][RS_LIST];

    const int4 item = p.items[blockIdx.x];
    const i32 first = item.y, end = min(item.z, p.n_slices);
    const i32 tx = item.x % p.ntx, tzy = item.x / p.ntx;
    const i32 ty = tzy % p.nty, tz = tzy / p.nty;
    const i32 wave = static_cast<i32>(threadIdx.x) >> 6, lane = static_cast<i32>(threadIdx.x) & 63;
    const i32 Z0 = p.base + tz * XT + wave * RS_Z, Y0 = p.base + ty * XT, X0 = tx * XT;   // this wave's slab
    const i32 n = p.n, width = n / 2 + 1;
    Out* vacc = vacc_all[wave];
    float* wacc = wacc_all[wave];
    unsigned char* tags = tags_all[wave];
    u32* wl = list[wave];

    for (i32 i = lane; i < RS_CELLS; i += 64) {
        vacc[i] = vzero<Out>();
        wacc[i] = 0.f;
    }
    lds_order();

    const float slice_f = static_cast<float>(n), target = static_cast<float>(n);
    const float M = p.margin;
    const float cz = static_cast<float>(Z0) + 0.5f * RS_Z, cy = static_cast<float>(Y0) + 0.5f * XT;
    const float cx = static_cast<float>(X0) + 0.5f * XT;
    for (i32 c0 = first; c0 < end; c0 += 64) {
        const bool cross = c0 + lane < end &&
                           slab_crossed<RS_Z>(plane_matrix(load_rotation(p.rot, c0 + lane),
                                                           slice_scaling(p.scale, p.scaled, c0 + lane)),
                                              cz, cy, cx, M);
        unsigned long long mask = __ballot(cross);
        while (mask) {
            const i32 k = __ffsll(static_cast<long long>(mask)) - 1;
            mask &= mask - 1;
            const i32 b = c0 + k;
            const Rot<float> r = load_rotation(p.rot, b);
            const Scale<float> sc = slice_scaling(p.scale, p.scaled, b);

            // fetch: noa's expressions up to the fractions and the input value and weight (the loads of two
            // candidates in flight per lane); store: the eight taps into LDS, and the x = 0 duplicates.
            struct Cand {
                Out value;
                float weight, tz, ty, tx;
                i32 lz, ly, lx;
                bool ok;
            };
            auto fetch = [&](i32 fy, i32 u, bool mirrored, Cand& c) {
                c.ok = false;
                vec2<float> f2;
                f2.y = static_cast<float>(fy) / slice_f;
                f2.x = static_cast<float>(u) / slice_f;
                // With a scaling: rasterize_kernel's general expression, without Ewald sphere (0 curvature).
                vec3<float> f = p.scaled ? slice2grid<true>(f2, sc, r, 0.f, 0.f) : slice2grid<false>(f2, sc, r, 0.f, 0.f);
                if (dot3(f) > p.cutoff_sqd) return;
                const bool flip = f.x < 0;
                if (flip != mirrored) return;
                RIn conjugate = 1;
                if (flip) {
                    f.z = -f.z; f.y = -f.y; f.x = -f.x;
                    if constexpr (vt<In>::cx) conjugate = -1;
                }
                const float fz = f.z * target, fyv = f.y * target, fx = f.x * target;
                const float flz = floorf(fz), fly = floorf(fyv), flx = floorf(fx);
                const i32 lz = static_cast<i32>(flz) - Z0, ly = static_cast<i32>(fly) - Y0;
                const i32 lx = static_cast<i32>(flx) - X0;
                if (static_cast<u32>(lz) >= static_cast<u32>(RS_Z) || static_cast<u32>(ly) >= static_cast<u32>(XT) ||
                    static_cast<u32>(lx) >= static_cast<u32>(XT))
                    return;
                const i32 y = frequency2index(p.slice_centered, fy, n);
                const In raw = p.in_is_value ? p.in_value : p.in[static_cast<i64>(b) * p.is_b + y * p.is_y + u * p.is_x];
                if constexpr (vt<In>::cx && CX) c.value = vcast<Out>(vmul(raw, conjugate));
                else c.value = vcast<Out>(raw);
                c.weight = 0.f;
                if (p.has_wvol) {
                    if (!p.win_present) c.weight = 1.f;
                    else c.weight = p.win_is_value ? p.win_value
                                                   : p.win[static_cast<i64>(b) * p.wis_b + y * p.wis_y + u * p.wis_x];
                }
                c.tz = fz - flz;
                c.ty = fyv - fly;
                c.tx = fx - flx;
                c.lz = lz; c.ly = ly; c.lx = lx;
                c.ok = true;
            };
            auto store = [&](const Cand& c) {
                // One lane per floor position does the plain read-modify-write, the others the atomic add.
                const i32 cell = (c.lz * XT + c.ly) * XT + c.lx;
                if (c.ok) tags[cell] = static_cast<unsigned char>(lane);
                lds_order();
                const bool own = c.ok && tags[cell] == static_cast<unsigned char>(lane);
                const bool shared = c.ok && !own;
                lds_order();
                float fr[2][3];
                fr[1][0] = c.tz; fr[1][1] = c.ty; fr[1][2] = c.tx;
                fr[0][0] = 1 - c.tz; fr[0][1] = 1 - c.ty; fr[0][2] = 1 - c.tx;
                const i32 o = (c.lz * XTT + c.ly) * XTT + c.lx;
#pragma unroll
                for (int w = 0; w < 2; ++w)
#pragma unroll
                    for (int v = 0; v < 2; ++v)
#pragma unroll
                        for (int uu = 0; uu < 2; ++uu) {
                            const float kk = fr[w][0] * fr[v][1] * fr[uu][2];
                            const i32 at = o + (w * XTT + v) * XTT + uu;
                            const Out add = vmul(c.value, kk);
                            const float wadd = c.weight * kk;
                            if (own) {
                                // Value and weight together (the weight is 0 without a weight volume).
                                const Out v0 = vacc[at];
                                const float w0 = wacc[at];
                                vacc[at] = vadd(v0, add);
                                wacc[at] = w0 + wadd;
                            }
                            lds_order();
                            if (shared) {
                                if constexpr (CX) {
                                    atomicAdd(&vacc[at].x, add.x);
                                    atomicAdd(&vacc[at].y, add.y);
                                } else {
                                    atomicAdd(vacc + at, add);
                                }
                                if (p.has_wvol) atomicAdd(wacc + at, wadd);
                            }
                            lds_order();
                        }
                // The x = 0 plane is redundant: the conjugate at (-w, -v, 0) with the same fraction (:347-372).
                if (c.ok && X0 + c.lx == 0) {
                    Out value = c.value;
                    if constexpr (CX) value.y = -value.y;
                    const i32 b0z = Z0 + c.lz, b0y = Y0 + c.ly;
#pragma unroll
                    for (int w = 0; w < 2; ++w) {
                        const i32 iw = frequency2index(p.vol_centered, -(b0z + w), n);
                        if (iw < 0 || iw >= n) continue;
#pragma unroll
                        for (int v = 0; v < 2; ++v) {
                            const i32 iv = frequency2index(p.vol_centered, -(b0y + v), n);
                            if (iv < 0 || iv >= n) continue;
                            const float kk = fr[w][0] * fr[v][1] * fr[0][2];
                            atomic_add(p.vol + (iw * p.vs0 + iv * p.vs1), vmul(value, kk));
                            if (p.has_wvol) atomic_add(p.wvol + (iw * p.wvs0 + iv * p.wvs1), c.weight * kk);
                        }
                    }
                }
            };

            tile_candidates<Cand, RS_Z, RS_LIST>(plane_matrix(r, sc), Z0, Y0, X0, n, M, p.pixels, p.inv_wr, p.slice_centered,
                                                 lane, wl, fetch, store);
        }
    }
    lds_order();

    // Flush: positions [Z0, Z0 + 4] x [Y0, Y0 + 16] x [X0, X0 + 16] at noa's indices, inside the stored volume.
    for (i32 i = lane; i < RS_CELLS; i += 64) {
        const i32 dz = i / (XTT * XTT), rem = i - dz * (XTT * XTT);
        const i32 dy = rem / XTT, dx = rem - dy * XTT;
        const i32 iu = X0 + dx;
        if (iu >= width) continue;
        const i32 iw = frequency2index(p.vol_centered, Z0 + dz, n);
        const i32 iv = frequency2index(p.vol_centered, Y0 + dy, n);
        if (iw < 0 || iw >= n || iv < 0 || iv >= n) continue;
        const Out v = vacc[i];
        if constexpr (CX) {
            if (v.x != 0.f || v.y != 0.f) atomic_add(p.vol + (iw * p.vs0 + iv * p.vs1 + iu * p.vs2), v);
        } else {
            if (v != 0.f) atomic_add(p.vol + (iw * p.vs0 + iv * p.vs1 + iu * p.vs2), v);
        }
        if (p.has_wvol) {
            const float w = wacc[i];
            if (w != 0.f) atomic_add(p.wvol + (iw * p.wvs0 + iv * p.wvs1 + iu * p.wvs2), w);
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

constexpr int NOT_TILED = -1000;

// The tiled path (rasterize_tile_kernel) where it applies: cubic volume and slices, no target shape, no Ewald
// sphere, cutoff <= 0.5, Mat33<f32> or Quaternion<f32> rotations, Mat22 inverse scalings or none. NOA_HIP_RASTERIZE_LAYOUT=direct|tiles forces either; by default
// the tiles take launches of at least 64 slices. NOA_HIP_RASTERIZE_PAIRS: expected crossing slices per work item.
template<class In, class Out, class C, class W>
int run_tiles(const ProjectDesc& d, const RasterParams<C, In, Out, W>& p, hipStream_t stream, std::string& err) {
    const i64 nn = p.grid[0], n_slices = d.in_shape[0];
    const bool cubic = p.grid[1] == nn && p.grid[2] == nn && d.in_shape[2] == nn && d.in_shape[3] == nn &&
                       d.target[0] == nn && d.target[1] == nn && d.target[2] == nn;
    const i64 stored = nn * nn * (nn / 2 + 1);
    // The candidate list of a wave must hold a whole slice row (n/2 + 1 pixels) for any matrix.
    const bool flat = p.ews[0] == 0 && p.ews[1] == 0;
    const bool eligible = cubic && flat && nn >= 16 && nn / 2 + 1 <= RS_LIST && d.cutoff <= 0.5 &&
                          (p.rot.kind == NOA_HIP_MAT33_F32 || p.rot.kind == NOA_HIP_QUAT_F32) &&
                          stored < 0x7fffffffLL && n_slices <= 0x7fffffffLL;
    if (!eligible) return NOT_TILED;
    const char* env = std::getenv("NOA_HIP_RASTERIZE_LAYOUT");
    const bool want = env ? std::strcmp(env, "tiles") == 0 : n_slices >= 64;
    if (!want) return NOT_TILED;
    int device = 0;
    (void)hipGetDevice(&device);
    i64 step = 1;
    while (n_slices > 16 * step) step *= 2;
    const i32 plan_slices = static_cast<i32>((n_slices + step - 1) / step * step);
    const char* pe = std::getenv("NOA_HIP_RASTERIZE_PAIRS");
    const i32 pairs = pe ? std::max(8, std::atoi(pe)) : -256;   // default: at most 256, fewer for small calls
    TilePlan plan;
    const int st = make_tile_plan(device, static_cast<i32>(nn), std::max(d.cutoff, 0.0), plan_slices, pairs, plan, err);
    if (st != NOA_HIP_OK) return st;
    if (plan.n_items == 0) return NOA_HIP_OK;
    RasterTileParams<In, Out> q{};
    q.in = p.in; q.is_b = p.is_b; q.is_y = p.is_y; q.is_x = p.is_x; q.in_is_value = p.in_is_value;
    q.in_value = p.in_value;
    q.win = p.win; q.wis_b = p.wis_b; q.wis_y = p.wis_y; q.wis_x = p.wis_x; q.win_present = p.win_present;
    q.win_is_value = p.win_is_value; q.win_value = p.win_value;
    q.vol = p.vol; q.vs0 = p.vs0; q.vs1 = p.vs1; q.vs2 = p.vs2;
    q.wvol = p.wvol; q.wvs0 = p.wvs0; q.wvs1 = p.wvs1; q.wvs2 = p.wvs2; q.has_wvol = p.has_wvol;
    q.rot = p.rot;
    q.scale = p.scale;
    q.scaled = p.general && p.scale.kind != NOA_HIP_XFORM_NONE;
    q.items = plan.items;
    q.n = static_cast<i32>(nn);
    q.nty = plan.nty; q.ntx = plan.ntx; q.base = plan.base;
    q.n_slices = static_cast<i32>(n_slices);
    q.pixels = p.pixels;
    q.inv_wr = 1.f / static_cast<float>(p.slice_wr);
    q.cutoff_sqd = p.cutoff_sqd;
    q.margin = 1e-3f + 1e-5f * static_cast<float>(nn);
    q.vol_centered = p.vol_centered;
    q.slice_centered = p.slices_centered;
    hipLaunchKernelGGL((rasterize_tile_kernel<In, Out>), dim3(static_cast<u32>(plan.n_items)), dim3(256), 0, stream,
                       q);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
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
    if constexpr (std::is_same_v<C, float> && std::is_same_v<W, float> &&
                  (std::is_same_v<In, float2> || std::is_same_v<In, float>) &&
                  (std::is_same_v<Out, float2> || std::is_same_v<Out, float>)) {
        const int st = run_tiles<In, Out>(d, p, stream, err);
        if (st != NOT_TILED) return st;
    }
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
