// Fourier insertion of central slices by 2D interpolation with a windowed sinc along the slice normal.
//
// Replaces FourierInsertInterpolate (src/noa/xform/FourierProject.hpp:392-539) launched by
// launch_insert_central_slices_3d (:1319-1394).
//
// noa computes, for every voxel of the stored half-volume, the sum over ALL slices (in slice order) of
// windowed_sinc(fz) * interp2(slice, f2), then adds it to the volume. Only slices whose slab |fz| <= f_b passes
// through the voxel contribute (about 0.6 % of the voxel-slice pairs at 256^3 with the default 1-voxel window);
// the others add exactly zero. This file keeps noa's per-voxel expressions and skips the zero terms.
//
// Culled paths (array inputs, single-precision values, no Ewald curvature):
//   - Work item = (8x8x16 sub-brick of the volume, range of slices), one wave per item, the sub-brick's
//     accumulator in LDS: no global atomics.
//   - The wave scans the item's slices 64 at a time: lane l tests whether the slab of slice s0+l meets the
//     sub-brick box; a ballot yields the passing slices, processed in increasing slice order.
//   - For a passing slice, the voxels near the slab are enumerated along the volume axis closest to the slab normal
//     and packed onto the wave's lanes; each candidate runs noa's exact per-voxel tests and accumulates in LDS. Each
//     voxel is met at most once per slice.
//   - Sub-bricks near the origin meet most slices; their slice range is cut into segments so that work items are
//     balanced. A sub-brick with one segment adds its sum to the volume directly; a split sub-brick writes one
//     partial sum per segment, and a second kernel adds the partials of each sub-brick in segment order. The
//     result is deterministic; it differs from noa's strict slice order only by this regrouping of the sum.
//   - insert_slab_kernel serves the common configuration (one power-of-two size, Mat33<f32> or Quaternion<f32>,
//     LINEAR, cutoff <= 0.5, no scaling) with per-slice records and a leaner per-voxel path; insert_wave_kernel
//     serves the others (inverse scalings, NEAREST, other sizes and cutoffs).
// General path (EWS, constant-value inputs, double precision): noa's per-voxel gather over every slice, one lane
// per voxel.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace noa_hip {

constexpr i32 NWIN = 12;               // degree of the window polynomial
constexpr i32 NT_REDUCE = 256;         // threads per workgroup of the partial-sum reduction

template<class C, class In, class W>
struct InsertParams {
    const In* in; i64 is_b; i32 is_y, is_x; bool in_is_value; In in_value;
    const W* win; i64 wis_b; i32 wis_y, wis_x; bool win_present, win_is_value; W win_value;
    void* vol; i32 vs0, vs1, vs2;
    W* wvol; i32 wvs0, wvs1, wvs2; bool has_wvol;
    i32 grid[3];          // logical D, H, W of the volume
    i32 wr;               // W/2+1
    i32 nb[3];            // bricks along z, y, x
    i32 bd[3];            // brick dims (voxels along z, y, x) of the launch
    i32 apitch;           // x pitch of a brick accumulator (and of its partial sums)
    const int4* items;    // {brick, first slice, end slice, partial slot or -1}
    const int4* reduce;   // {brick, first partial slot, number of partials, 0}
    float* partials;      // per slot: ACC complex accumulators then ACC weights (as float pairs / floats)
    C target[3];
    C slice_f[2];
    i32 slice_shape[2];   // logical H, W of the slices
    i32 n_slices, slice_begin, slice_end;
    C cutoff_sqd, fftfreq_sinc, fftfreq_blackman;
    C ews[2];
    XformArg<C> rot, scale;
    bool general, slices_centered, vol_centered;
    C inv_target[3];      // 1 / target
    bool pow2;            // every target is a power of two (then x / t == x * (1 / t) exactly)
    float wa[NWIN + 1];   // windowed sinc as a polynomial in t = 2 (fz/f_b)^2 - 1 (single-precision culled paths)
    bool wpoly;           // use wa; otherwise evaluate noa's expression
    float inv_fb;         // 1 / fftfreq_blackman
    u32 slice_bytes, wslice_bytes;   // byte extent of one slice and of its weights (buffer loads)
    int interp_mode;      // the method of the INTERP_WINDOW instantiation (CUBIC, CUBIC_BSPLINE, LANCZOS*)
};

// Keeps a value in vector registers: the compiler would otherwise re-read kernel arguments inside the loops,
// and each scalar load there makes the wave wait for all of its outstanding LDS operations.
template<class T> __device__ __forceinline__ T pin(T v) {
    asm volatile("" : "+v"(v));
    return v;
}

// Frequency of index `i` along a volume axis; `is_x`: the rfft half axis (index == frequency).
__device__ __forceinline__ i32 axis_freq(bool is_x, i32 i, bool centered, i32 n) {
    return is_x ? i : index2frequency(centered, i, n);
}

// Windowed sinc of the culled kernels in single precision: a polynomial of degree NWIN in t = 2 (fz/f_b)^2 - 1
// (the window is even in fz), fitted on the host per (f_s, f_b) and accepted only when its error, evaluated
// in single precision, stays below 4e-7 on [-f_b, f_b] (noa's FP32 expression itself carries ~2e-7 of rounding).
__device__ __forceinline__ float window_poly(float fz, float inv_fb, const float (&a)[NWIN + 1]) {
    const float q = fz * inv_fb;
    const float t = fmaf(2.f * q, q, -1.f);
    float r = a[NWIN];
#pragma unroll
    for (i32 k = NWIN - 1; k >= 0; --k) r = fmaf(r, t, a[k]);
    return r;
}



// ---------------------------------------------------------------------------------------------------------
// One-wave path: a work item is (8x8x16 sub-brick, range of slices) processed by one wave with no barrier
// between slices. The accumulator (12 KiB of LDS) is private to the wave; the slice taps are read from global
// memory (L2-resident footprints). Each voxel is owned by one lane per slice and slices are processed in order,
// so the sums are deterministic.
// ---------------------------------------------------------------------------------------------------------
constexpr i32 WZ = 8, WY = 8, WX = 16;     // sub-brick voxels along z, y, x (WZ == WY: the face axis B has 8)
constexpr i32 WVOX = WZ * WY * WX;

// Enumeration data of a (sub-brick, slice) pair: the volume axis A closest to the slab normal (in index units)
// and the normal's components along A (inverted) and the face axes B < Cc.
template<class C>
struct AxisInfo {
    i32 flags;            // bits 0-1: axis A, 4: candidate range from the slab
    C inv_alA, alB, alC;
};

template<class C>
__device__ __forceinline__ AxisInfo<C> axis_info(const C m[9], C it0, C it1, C it2, const bool contiguous[3]) {
    AxisInfo<C> r;
    const C al0 = m[0] * it0, al1 = m[1] * it1, al2 = m[2] * it2;
    i32 ax = 0;
    if (fabs(al1) > fabs(al0)) ax = 1;
    if (fabs(al2) > fabs(ax == 0 ? al0 : al1)) ax = 2;
    const C alA = ax == 0 ? al0 : (ax == 1 ? al1 : al2);
    r.alB = ax == 0 ? al1 : al0;
    r.alC = ax == 2 ? al1 : al2;
    const bool contig = ax == 0 ? contiguous[0] : (ax == 1 ? contiguous[1] : contiguous[2]);
    const bool ranged = contig && fabs(alA) > C(1e-20);
    r.inv_alA = ranged ? C(1) / alA : C(0);
    r.flags = ax | (ranged ? 16 : 0);
    return r;
}

constexpr i32 WG = 2;                      // candidates per lane and pass

// The taps of one 2D interpolation of a slice and of its weights (Interpolation.hpp:435-526): noa's flip of the
// frequency when x < 0, the fractions, and the four (LINEAR) or one (NEAREST) tap values, each loaded from a
// clamped address and zeroed when outside the frequency bounds. All loads are issued before any is used.
template<class In, class W, class C>
struct Taps {
    In v[2][2];           // raw loads; zeroed in combine_taps where `ok` is clear (loads stay in flight until then)
    W w[2][2];
    u32 ok;               // bit (y*2 + x): tap inside the bounds
    typename vt<In>::real ty, tx, conj;
};

// A slice (or its weights) read through a buffer resource: base and byte size in scalar registers, a 32-bit byte
// offset per load, and out-of-range offsets read 0.
struct SliceBuf {
    __amdgpu_buffer_rsrc_t rsrc;
    __device__ __forceinline__ SliceBuf(const void* base, u32 bytes)
        : rsrc(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000)) {}
    template<class T> __device__ __forceinline__ T load(u32 elem) const {
        return load_bytes<T>(elem * static_cast<u32>(sizeof(T)));
    }
    template<class T> __device__ __forceinline__ T load_bytes(u32 off) const {
        T v;
        if constexpr (sizeof(T) == 4) {
            const u32 r = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0);
            __builtin_memcpy(&v, &r, 4);
        } else if constexpr (sizeof(T) == 8) {
            const auto r = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0);
            __builtin_memcpy(&v, &r, 8);
        } else {
            const auto r = __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0);
            __builtin_memcpy(&v, &r, 16);
        }
        return v;
    }
};

// Element index past the end of any slice or weight buffer (byte offsets 0xFFFFFFF8 for 8-byte values,
// 0x7FFFFFFC for 4-byte weights): a buffer load there returns 0.
constexpr u32 TAP_OUTSIDE = 0x1FFFFFFFu;

template<int INTERP, bool ROWS, class In, class W, class C>
__device__ __forceinline__ Taps<In, W, C> fetch_taps(const SliceBuf& vb, i32 vsy, i32 vsx, const SliceBuf& wb,
                                                     i32 wsy, i32 wsx, bool has_w, i32 ny, i32 nx,
                                                     bool centered, C fy, C fx, bool valid) {
    Taps<In, W, C> t;
    t.conj = 1;
    if (fx < 0) { fy *= -1; fx *= -1; t.conj = -1; }
    const i32 lo_y = -ny / 2, hi_y = (ny - 1) / 2, hi_x = nx / 2;
    if constexpr (ROWS && INTERP == NOA_HIP_INTERP_LINEAR) {
        // Contiguous rows shared by the values and the weights (unit x strides, equal row pitches): one row offset
        // per tap row, and a tap outside the bounds reads TAP_OUTSIDE, which the buffer load returns as 0 -- the
        // zero noa's interpolation adds for it.
        const C fly = floor(fy), flx = floor(fx);
        const i32 y0 = static_cast<i32>(fly), x0 = static_cast<i32>(flx);
        t.ty = static_cast<typename vt<In>::real>(fy - fly);
        t.tx = static_cast<typename vt<In>::real>(fx - flx);
        const bool okx0 = x0 >= 0 && x0 <= hi_x, okx1 = x0 + 1 >= 0 && x0 + 1 <= hi_x;
        t.ok = 0xfu;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const bool oky = valid && y0 + y >= lo_y && y0 + y <= hi_y;
            const i32 row = frequency2index(centered, y0 + y, ny) * vsy;
            const u32 e0 = oky && okx0 ? static_cast<u32>(row + x0) : TAP_OUTSIDE;
            const u32 e1 = oky && okx1 ? static_cast<u32>(row + x0 + 1) : TAP_OUTSIDE;
            t.v[y][0] = vb.template load<In>(e0);
            t.v[y][1] = vb.template load<In>(e1);
            t.w[y][0] = wb.template load<W>(e0);
            t.w[y][1] = wb.template load<W>(e1);
        }
        (void)vsx; (void)wsy; (void)wsx; (void)has_w;
    } else if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        const i32 y = static_cast<i32>(round(fy)), x = static_cast<i32>(round(fx));
        const bool ok = valid && y >= lo_y && y <= hi_y && x >= 0 && x <= hi_x;
        const i32 iy = ok ? frequency2index(centered, y, ny) : 0, ix = ok ? x : 0;
        t.v[0][0] = vb.template load<In>(static_cast<u32>(iy * vsy + ix * vsx));
        t.w[0][0] = has_w ? wb.template load<W>(static_cast<u32>(iy * wsy + ix * wsx)) : W(0);
        t.ok = ok ? 1u : 0u;
        t.ty = 0;
        t.tx = 0;
    } else {
        const C fly = floor(fy), flx = floor(fx);
        const i32 y0 = static_cast<i32>(fly), x0 = static_cast<i32>(flx);
        t.ty = static_cast<typename vt<In>::real>(fy - fly);
        t.tx = static_cast<typename vt<In>::real>(fx - flx);
        const bool okx0 = x0 >= 0 && x0 <= hi_x, okx1 = x0 + 1 >= 0 && x0 + 1 <= hi_x;
        t.ok = 0;
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const bool oky = valid && y0 + y >= lo_y && y0 + y <= hi_y;
            const i32 iy = oky ? frequency2index(centered, y0 + y, ny) : 0;
            const i32 x = oky && okx0 ? x0 : 0;
            t.ok |= (oky && okx0 ? 1u : 0u) << (y * 2);
            t.ok |= (oky && okx1 ? 1u : 0u) << (y * 2 + 1);
            // One load per tap: 16-byte loads of both x taps were measured slower (misaligned at odd x).
            const i32 x1 = oky && okx1 ? x0 + 1 : 0;
            t.v[y][0] = vb.template load<In>(static_cast<u32>(iy * vsy + x * vsx));
            t.v[y][1] = vb.template load<In>(static_cast<u32>(iy * vsy + x1 * vsx));
            t.w[y][0] = has_w ? wb.template load<W>(static_cast<u32>(iy * wsy + x * wsx)) : W(0);
            t.w[y][1] = has_w ? wb.template load<W>(static_cast<u32>(iy * wsy + x1 * wsx)) : W(0);
        }
    }
    return t;
}

// Sums the taps in noa's order: x innermost, then y; then the imaginary part is negated if the frequency was
// flipped (LINEAR only; NEAREST does not conjugate, Interpolation.hpp:479-484).
template<int INTERP, class In, class W, class C>
__device__ __forceinline__ void combine_taps(const Taps<In, W, C>& t, bool has_w, In& value, W& weight) {
    using R = typename vt<In>::real;
    if constexpr (INTERP == NOA_HIP_INTERP_NEAREST) {
        value = vadd(vzero<In>(), vmul(t.ok ? t.v[0][0] : vzero<In>(), R(1)));
        weight = W(0) + (t.ok ? t.w[0][0] : W(0)) * W(1);
    } else {
        const R wy[2] = {1 - t.ty, t.ty}, wx[2] = {1 - t.tx, t.tx};
        value = vzero<In>();
        weight = W(0);
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            In vy = vzero<In>();
            W wyv = W(0);
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                const bool ok = (t.ok >> (y * 2 + x)) & 1u;
                vy = vadd(vy, vmul(ok ? t.v[y][x] : vzero<In>(), wx[x]));
                wyv = wyv + (ok ? t.w[y][x] : W(0)) * static_cast<W>(wx[x]);
            }
            value = vadd(value, vmul(vy, wy[y]));
            weight = weight + wyv * static_cast<W>(wy[y]);
        }
        vimag_mul(value, t.conj);
    }
    (void)has_w;
}

// HOT: 0 = every option read at run time; 1 and 2 = the common configuration with its options fixed at compile
// time (power-of-two target, Mat33 rotations, input and output weights with unit x strides and one row pitch,
// polynomial window; 1: h2h, 2: hc2hc), which frees scalar registers, removes the uniform branches of the hot loop
// and shares one row offset between the value and weight taps (same results).
template<class C, class In, class Out, class W, int INTERP, int HOT = 0>
__device__ __forceinline__ void insert_wave_item(const InsertParams<C, In, W>& p, int4 item, In* __restrict__ acc_v,
                                                 W* __restrict__ acc_w, i32* __restrict__ cand) {
    using RIn = typename vt<In>::real;
    const i32 lane = static_cast<i32>(threadIdx.x);
    // The plan is built for a rounded-up slice count: clamp the item's slice range to the actual count (an empty
    // range still flushes, so a split brick's partial slot is always written).
    item.z = min(item.z, p.n_slices);
    item.y = min(item.y, item.z);
    const i32 brick = item.x;
    const i32 bxi = brick % p.nb[2];
    const i32 rem = brick / p.nb[2];
    const i32 byi = rem % p.nb[1];
    const i32 bzi = rem / p.nb[1];
    constexpr i32 dims[3] = {WZ, WY, WX};
    const i32 org[3] = {bzi * WZ, byi * WY, bxi * WX};
    const i32 grid[3] = {p.grid[0], p.grid[1], p.grid[2]};
    const i32 ext[3] = {p.grid[0], p.grid[1], p.wr};
    const bool vc = HOT == 0 ? p.vol_centered : HOT == 2;

    for (i32 v = lane; v < WVOX; v += 64) {
        acc_v[v] = vzero<In>();
        acc_w[v] = W(0);
    }

    // Sub-brick box in normalised frequencies, and whether each axis is free of the non-centred wrap (then the
    // frequencies along it are first-frequency + k).
    C cen[3], half[3];
    i32 f0[3], n_in[3];
    bool contiguous[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        i32 lo = 0x7fffffff, hi = -0x7fffffff, n = 0;
        bool contig = true;
        const i32 first = axis_freq(a == 2, org[a], vc, grid[a]);
        for (i32 k = 0; k < dims[a]; ++k) {
            const i32 idx = org[a] + k;
            if (idx >= ext[a]) break;
            const i32 f = axis_freq(a == 2, idx, vc, grid[a]);
            contig = contig && (f == first + k);
            lo = min(lo, f);
            hi = max(hi, f);
            ++n;
        }
        f0[a] = first;
        n_in[a] = n;
        contiguous[a] = contig;
        cen[a] = (static_cast<C>(lo) + static_cast<C>(hi)) * C(0.5) / p.target[a];
        half[a] = (static_cast<C>(hi) - static_cast<C>(lo)) * C(0.5) / p.target[a];
    }
    const C slack = C(1e-5);
    const i32 Hs = p.slice_shape[0], Ws = p.slice_shape[1];
    const bool all_contig = contiguous[0] && contiguous[1] && contiguous[2];

    // Values of the hot loop, held in vector registers.
    const C fb = pin(p.fftfreq_blackman);
    const C fs = pin(p.fftfreq_sinc);
    const C cutoff_sqd = pin(p.cutoff_sqd);
    const C hs_f = pin(p.slice_f[0]), ws_f = pin(p.slice_f[1]);
    const C t0 = pin(p.target[0]), t1 = pin(p.target[1]), t2 = pin(p.target[2]);
    const C it0 = pin(p.inv_target[0]), it1 = pin(p.inv_target[1]), it2 = pin(p.inv_target[2]);
    const bool pow2 = HOT != 0 || p.pow2;
    const bool wpoly = HOT != 0 || p.wpoly;
    const bool has_wvol = HOT != 0 || p.has_wvol;
    const bool win_present = HOT != 0 || p.win_present;
    float wa[NWIN + 1];
#pragma unroll
    for (int k = 0; k <= NWIN; ++k) wa[k] = pin(p.wa[k]);
    const float inv_fb = pin(p.inv_fb);

    // Iterator over the passing slices of the item, in increasing order: lane l holds slice chunk0+l, tests its
    // slab against the sub-brick box and prepares the enumeration; a passing slice's values are broadcast from
    // that lane. The next chunk's rotations are loaded one chunk ahead.
    const bool quat = HOT == 0 && (p.rot.kind == NOA_HIP_QUAT_F32 || p.rot.kind == NOA_HIP_QUAT_F64);
    auto scan_load = [&](i32 s0, Rot<C>& r) {
        const i32 s = s0 + lane;
        if (s < item.z) r = load_rotation(p.rot, s);
    };
    i32 chunk0 = item.y;
    Rot<C> chunk{}, ahead{};
    chunk.quat = quat;
    ahead.quat = quat;
    AxisInfo<C> mine{};
    unsigned long long mask = 0;
    auto scan_prepare = [&]() {
        bool pass = false;
        if (chunk0 + lane < item.z) {
            C m[9];
            rot_matrix(chunk, m);
            const C d = fabs(m[0] * cen[0] + m[1] * cen[1] + m[2] * cen[2]);
            const C bound = fb + fabs(m[0]) * half[0] + fabs(m[1]) * half[1] + fabs(m[2]) * half[2];
            pass = d <= bound + slack;
            mine = axis_info(m, it0, it1, it2, contiguous);
        }
        mask = __ballot(pass);
    };
    scan_load(chunk0, chunk);
    scan_load(chunk0 + 64, ahead);
    scan_prepare();
    auto next_slice = [&](Rot<C>& rot_out, AxisInfo<C>& out) -> i32 {
        while (mask == 0) {
            chunk0 += 64;
            if (chunk0 >= item.z) return -1;
#pragma unroll
            for (int q = 0; q < 9; ++q) chunk.m[q] = ahead.m[q];
            scan_load(chunk0 + 64, ahead);
            scan_prepare();
        }
        const i32 k = __ffsll(static_cast<long long>(mask)) - 1;
        mask &= mask - 1;
        rot_out.quat = quat;
#pragma unroll
        for (int q = 0; q < 9; ++q) rot_out.m[q] = bcast(chunk.m[q], k);
        out.flags = __builtin_amdgcn_readlane(mine.flags, k);
        out.inv_alA = bcast(mine.inv_alA, k);
        out.alB = bcast(mine.alB, k);
        out.alC = bcast(mine.alC, k);
        return chunk0 + k;
    };

    // Inverse scalings (no Ewald curvature on this path): the 2D frequency of grid2slice is scaled, fz is not.
    const bool scaled = HOT == 0 && p.general;
    Rot<C> rot{};
    AxisInfo<C> ai{};
    for (i32 cur = next_slice(rot, ai); cur >= 0; cur = next_slice(rot, ai)) {
        Scale<C> sc;
        sc.identity = true;
        if (scaled) sc = load_scaling(p.scale, cur);
        // Face of axis A: B (z, or y when A is z; 8 voxels) and Cc (x, or y when A is x). Columns: 8 x FC.
        const i32 ax = ai.flags & 3;
        const bool ranged = ai.flags & 16;
        const i32 lc_bits = ax == 2 ? 3 : 4;
        const i32 fc = 1 << lc_bits;
        const i32 n_col = WZ << lc_bits;
        const i32 orgB = ax == 0 ? org[1] : org[0];
        const i32 orgC = ax == 2 ? org[1] : org[2];
        const i32 gridB = ax == 0 ? grid[1] : grid[0];
        const i32 gridC = grid[1];                                  // unused when Cc is x
        const bool xC = ax != 2;
        const i32 extB = ax == 0 ? ext[1] : ext[0];
        const i32 extC = ax == 2 ? ext[1] : ext[2];
        const i32 f0A = ax == 0 ? f0[0] : (ax == 1 ? f0[1] : f0[2]);
        const i32 nA = ax == 0 ? n_in[0] : (ax == 1 ? n_in[1] : n_in[2]);
        const C tA = ax == 0 ? t0 : (ax == 1 ? t1 : t2);

        // Candidates of the lane's columns (lane and lane + 64): |alA * f + beta| <= f_b in exact arithmetic,
        // widened by eps samples to cover the rounding of the exact per-voxel test (a few ulps of |f| <= 1,
        // divided by |alA| >= 1/(sqrt(3) T)).
        // With contiguous face axes the range centre is linear in the column: centre = u0 + ub lb + uc lc (in
        // samples along A, relative to the sub-brick's first sample), computed with two FMAs.
        const bool contigB = ax == 0 ? contiguous[1] : contiguous[0];
        const bool contigC = ax == 2 ? contiguous[1] : contiguous[2];
        const bool linear = ranged && contigB && contigC;
        const i32 f0B = ax == 0 ? f0[1] : f0[0];
        const i32 f0C = ax == 2 ? f0[1] : f0[2];
        const C eps_l = C(1e-3) + C(8e-6) * tA;
        const C ub = -ai.alB * ai.inv_alA, uc = -ai.alC * ai.inv_alA;
        const C u0 = -(ai.alB * static_cast<C>(f0B) + ai.alC * static_cast<C>(f0C)) * ai.inv_alA - static_cast<C>(f0A);
        const C halfw = fabs(fb * ai.inv_alA) + eps_l;
        i32 cnt[2] = {0, 0}, klo[2] = {0, 0};
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const i32 col = lane + 64 * j;
            if (col >= n_col) continue;
            const i32 lb = col >> lc_bits, lc = col & (fc - 1);
            const i32 gb = orgB + lb, gc = orgC + lc;
            if (gb >= extB || gc >= extC) continue;
            i32 k_lo = 0, k_hi = nA - 1;
            if (linear) {
                const C centre = fma(ub, static_cast<C>(lb), fma(uc, static_cast<C>(lc), u0));
                const C kl = ceil(centre - halfw), kh = floor(centre + halfw);
                k_lo = static_cast<i32>(fmin(fmax(kl, C(0)), static_cast<C>(nA)));
                k_hi = static_cast<i32>(fmax(fmin(kh, static_cast<C>(nA - 1)), C(-1)));
            } else if (ranged) {
                const i32 fbq = index2frequency(vc, gb, gridB);
                const i32 fcq = axis_freq(xC, gc, vc, gridC);
                const C beta = ai.alB * static_cast<C>(fbq) + ai.alC * static_cast<C>(fcq);
                const C eps = C(1e-3) + C(8e-6) * tA;
                C lo = (-fb - beta) * ai.inv_alA;
                C hi = (fb - beta) * ai.inv_alA;
                if (lo > hi) { const C t = lo; lo = hi; hi = t; }
                const C kl = ceil(lo - eps) - static_cast<C>(f0A);
                const C kh = floor(hi + eps) - static_cast<C>(f0A);
                k_lo = static_cast<i32>(fmin(fmax(kl, C(0)), static_cast<C>(nA)));
                k_hi = static_cast<i32>(fmax(fmin(kh, static_cast<C>(nA - 1)), C(-1)));
            }
            cnt[j] = max(0, k_hi - k_lo + 1);
            klo[j] = k_lo;
        }
        const i32 mine_cnt = cnt[0] + cnt[1];
        const i32 incl = wave_inclusive_sum(mine_cnt);
        const i32 total = __builtin_amdgcn_readlane(incl, 63);
        const i32 excl = incl - mine_cnt;

        const SliceBuf sl(p.in + static_cast<i64>(cur) * p.is_b, p.slice_bytes);
        const SliceBuf wsl(win_present ? static_cast<const void*>(p.win + static_cast<i64>(cur) * p.wis_b)
                                       : static_cast<const void*>(p.in), win_present ? p.wslice_bytes : 0u);
        for (i32 base = 0; base < total; base += 64 * WG) {
            // Pack the candidates [base, base + 64 WG) of the wave into the list, WG per lane.
            __syncthreads();
            {
                const i32 i0 = max(0, base - excl), i1 = min(mine_cnt, base + 64 * WG - excl);
                for (i32 i = i0; i < i1; ++i) {
                    const bool second = i >= cnt[0];
                    const i32 col = second ? lane + 64 : lane;
                    const i32 k = second ? klo[1] + (i - cnt[0]) : klo[0] + i;
                    cand[excl + i - base] = (col << 4) | k;
                }
            }
            __syncthreads();
            // Phase 1: noa's exact per-voxel tests (:479-494), the window, and every tap load of the WG candidates
            // issued together.
            Taps<In, W, C> tp[WG];
            C window[WG];
            i32 lv[WG];
            bool valid[WG];
#pragma unroll
            for (int g = 0; g < WG; ++g) {
                valid[g] = false;
                if (base + g * 64 >= total) continue;   // uniform: an empty group of 64 is skipped
                const i32 e = cand[g * 64 + lane];
                const i32 k = e & (ax == 2 ? 15 : 7);
                const i32 col = (e >> 4) & (n_col - 1);
                const i32 lb = col >> lc_bits, lc = col & (fc - 1);
                const i32 lz = ax == 0 ? k : lb;
                const i32 ly = ax == 1 ? k : (ax == 0 ? lb : lc);
                const i32 lx = ax == 2 ? k : lc;
                const i32 frz = all_contig ? f0[0] + lz : index2frequency(vc, org[0] + lz, grid[0]);
                const i32 fry = all_contig ? f0[1] + ly : index2frequency(vc, org[1] + ly, grid[1]);
                const i32 frx = org[2] + lx;
                // The normalised frequency as noa computes it: index2frequency / target (:479-484).
                vec3<C> f;
                if (pow2) {
                    f.z = static_cast<C>(frz) * it0; f.y = static_cast<C>(fry) * it1; f.x = static_cast<C>(frx) * it2;
                } else {
                    f.z = static_cast<C>(frz) / t0; f.y = static_cast<C>(fry) / t1; f.x = static_cast<C>(frx) / t2;
                }
                vec2<C> f2;
                const C fz = scaled ? grid2slice<true>(f, sc, rot, C(0), C(0), f2)
                                    : grid2slice<false>(f, sc, rot, C(0), C(0), f2);
                valid[g] = base + g * 64 + lane < total && !(dot3(f) > cutoff_sqd) && fabs(fz) <= fb;
                if constexpr (sizeof(C) == 4) {
                    window[g] = wpoly ? window_poly(fz, inv_fb, wa) : windowed_sinc<C>(fz, fs, fb);
                } else {
                    window[g] = windowed_sinc<C>(fz, fs, fb);
                }
                lv[g] = (lz * WY + ly) * WX + lx;
                tp[g] = fetch_taps<INTERP, HOT != 0, In, W>(sl, p.is_y, p.is_x, wsl, p.wis_y, p.wis_x, win_present,
                                                     Hs, Ws, HOT == 0 ? p.slices_centered : HOT == 2, f2.y * hs_f,
                                                     f2.x * ws_f, valid[g]);
            }
            // Phase 2: interpolate (Interpolation.hpp:435-526) and accumulate.
#pragma unroll
            for (int g = 0; g < WG; ++g) {
                if (base + g * 64 >= total || !valid[g]) continue;
                In v;
                W wv;
                combine_taps<INTERP>(tp[g], win_present, v, wv);
                acc_v[lv[g]] = vadd(acc_v[lv[g]], vmul(v, static_cast<RIn>(window[g])));
                if (has_wvol)
                    acc_w[lv[g]] = acc_w[lv[g]] + (win_present ? wv * static_cast<W>(window[g]) : static_cast<W>(window[g]));
            }
        }
    }
    __syncthreads();

    if (item.w >= 0) {
        // Split sub-brick: store this segment's sums (every slot; the reduction applies the cutoff).
        In* pv = reinterpret_cast<In*>(p.partials + static_cast<i64>(item.w) * WVOX * 3);
        W* pw = reinterpret_cast<W*>(p.partials + static_cast<i64>(item.w) * WVOX * 3 + WVOX * 2);
        for (i32 v = lane; v < WVOX; v += 64) {
            pv[v] = acc_v[v];
            pw[v] = acc_w[v];
        }
        return;
    }
    // value and weights of every in-cutoff voxel are added to the volume (:516-518).
    for (i32 v = lane; v < WVOX; v += 64) {
        const i32 lz = v / (WY * WX), ly = (v / WX) % WY, lx = v % WX;
        const i32 z = org[0] + lz, y = org[1] + ly, x = org[2] + lx;
        if (z >= ext[0] || y >= ext[1] || x >= ext[2]) continue;
        vec3<C> f;
        f.z = static_cast<C>(index2frequency(vc, z, grid[0])) / p.target[0];
        f.y = static_cast<C>(index2frequency(vc, y, grid[1])) / p.target[1];
        f.x = static_cast<C>(x) / p.target[2];
        if (dot3(f) > p.cutoff_sqd) continue;
        Out* o = static_cast<Out*>(p.vol) + (z * p.vs0 + y * p.vs1 + x * p.vs2);
        *o = vadd(*o, vcast<Out>(acc_v[v]));
        if (p.has_wvol) {
            W* w = p.wvol + (z * p.wvs0 + y * p.wvs1 + x * p.wvs2);
            *w = *w + acc_w[v];
        }
    }
}

template<class C, class In, class Out, class W, int INTERP, int HOT = 0>
__global__ __launch_bounds__(64) void insert_wave_kernel(const InsertParams<C, In, W> p) {
    __shared__ In acc_v[WVOX];
    __shared__ W acc_w[WVOX];
    __shared__ i32 cand[64 * WG];          // packed candidates of one pass: column << 4 | k
    const int4 item = p.items[blockIdx.x];
    if (item.x < 0) return;                 // padding of an XCD-ordered plan
    insert_wave_item<C, In, Out, W, INTERP, HOT>(p, item, acc_v, acc_w, cand);
}

// Batched insertion: several insertion calls (each its own volume, box size, layout, slice count and cutoff) in one
// persistent launch of the one-wave kernel. The device image of a batch holds a header, one parameter block per call
// and the work items of every call ({sub-brick, first slice, end slice, call}; the calls' slice ranges are not split,
// so no partial sums); workgroup g takes the items g, g + G, ... (items of each call sorted by expected work, calls
// interleaved by the host). Each item runs the one-wave kernel's code with its call's parameters: the results are
// those of the calls made one by one (same per-voxel sums, in slice order).
struct InsertBatchHeader {
    i32 n_items, n_jobs;
    i32 jobs_offset, items_offset;       // bytes from the start of the image
    i32 pad[12];
};
static_assert(sizeof(InsertBatchHeader) == 64, "batch header");

__global__ __launch_bounds__(64) void insert_batch_kernel(const unsigned char* __restrict__ image) {
    using P = InsertParams<float, float2, float>;
    __shared__ float2 acc_v[WVOX];
    __shared__ float acc_w[WVOX];
    __shared__ i32 cand[64 * WG];
    const InsertBatchHeader* h = reinterpret_cast<const InsertBatchHeader*>(image);
    const P* __restrict__ jobs = reinterpret_cast<const P*>(image + h->jobs_offset);
    const int4* __restrict__ items = reinterpret_cast<const int4*>(image + h->items_offset);
    const i32 n_items = h->n_items;
    for (i32 i = static_cast<i32>(blockIdx.x); i < n_items; i += static_cast<i32>(gridDim.x)) {
        int4 item = items[i];
        const P& p = jobs[item.w];
        item.w = -1;
        __syncthreads();   // the previous item's flush has read the accumulator
        insert_wave_item<float, float2, float2, float, NOA_HIP_INTERP_LINEAR, 0>(p, item, acc_v, acc_w, cand);
    }
}

// Adds the partial sums of each split brick to the volume, in segment order.
template<class C, class In, class Out, class W>
__global__ __launch_bounds__(NT_REDUCE) void insert_reduce_kernel(const InsertParams<C, In, W> p) {
    const int4 red = p.reduce[blockIdx.x];
    const i32 brick = red.x;
    const i32 bxi = brick % p.nb[2];
    const i32 rem = brick / p.nb[2];
    const i32 bd0 = p.bd[0], bd1 = p.bd[1], bd2 = p.bd[2];
    const i32 org[3] = {(rem / p.nb[1]) * bd0, (rem % p.nb[1]) * bd1, bxi * bd2};
    const i32 ext[3] = {p.grid[0], p.grid[1], p.wr};
    const i64 slot = static_cast<i64>(bd0) * bd1 * p.apitch;
    for (i32 v = static_cast<i32>(threadIdx.x); v < bd0 * bd1 * bd2; v += NT_REDUCE) {
        const i32 lz = v / (bd1 * bd2), ly = (v / bd2) % bd1, lx = v % bd2;
        const i32 z = org[0] + lz, y = org[1] + ly, x = org[2] + lx;
        if (z >= ext[0] || y >= ext[1] || x >= ext[2]) continue;
        vec3<C> f;
        f.z = static_cast<C>(index2frequency(p.vol_centered, z, p.grid[0])) / p.target[0];
        f.y = static_cast<C>(index2frequency(p.vol_centered, y, p.grid[1])) / p.target[1];
        f.x = static_cast<C>(x) / p.target[2];
        if (dot3(f) > p.cutoff_sqd) continue;
        const i32 a = (lz * bd1 + ly) * p.apitch + lx;
        In sv = vzero<In>();
        W sw = W(0);
        for (i32 s = 0; s < red.z; ++s) {
            const float* base = p.partials + static_cast<i64>(red.y + s) * slot * 3;
            sv = vadd(sv, reinterpret_cast<const In*>(base)[a]);
            sw = sw + reinterpret_cast<const W*>(base + slot * 2)[a];
        }
        Out* o = static_cast<Out*>(p.vol) + (z * p.vs0 + y * p.vs1 + x * p.vs2);
        *o = vadd(*o, vcast<Out>(sv));
        if (p.has_wvol) {
            W* w = p.wvol + (z * p.wvs0 + y * p.wvs1 + x * p.wvs2);
            *w = *w + sw;
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Slab kernel: the common single-precision configuration (Mat33<f32> rotations, c32 or f32 slices with f32 or no
// weights, LINEAR, one power-of-two size n for the slices, the volume and the target, cutoff <= 0.5, no scaling
// or Ewald sphere). Same work items and sub-bricks as the one-wave kernel, with less work per pair and per voxel:
//   - A prologue kernel writes one record per slice: the rotation, the volume axis A closest to the slab normal
//     and the slab's centre and half-width along A as linear functions of the two other coordinates.
//   - Per (sub-brick, slice) pair, each lane takes one or two columns of the sub-brick face across A, clamps the
//     slab range of its column to the sub-brick and to the cutoff sphere (exact integer test, precomputed per
//     column), and the candidates are packed into an LDS list by ballots (one per candidate rank j).
//   - Per candidate, the coordinates are noa's expressions in sample units: with n a power of two,
//     R * (index / n) * n rounds exactly like R * index, so pz, py, px are noa's fz*n, f2y*n, f2x*n bit for bit.
//     The window is a degree-9 polynomial in t = 2 (pz/fbT)^2 - 1 (fitted per window, error < 4e-7), folded
//     into the bilinear weights; taps beyond noa's frequency bounds get weight 0 and never load outside the slice
//     (buffer resources). Sums are added to the LDS accumulator (plain read-modify-write: a voxel is met at most
//     once per slice and slices are processed in order, so the result is deterministic). The tap loads of the next
//     passing slice are issued before the current slice's sums are added.
// ---------------------------------------------------------------------------------------------------------
constexpr i32 SLAB_NWIN = 9;        // degree of the window polynomial
// Candidates of one pair: at most 4 per column (the slab kernel runs for fbT <= 1.15 samples, so a column's slab
// range is shorter than 2 * (1.15 * sqrt(3) + 2e-3) < 4 samples) over at most 128 columns (the 8 x 16 faces).
constexpr i32 SLAB_MAXC = 512;
static_assert(WZ == 8 && WY == 8 && WX == 16, "slab candidate words decode lz = v >> 7, ly = (v >> 4) & 7, lx = v & 15");

template<class In, class W>
struct SlabParams {
    const In* in; i64 is_b;             // slices, unit x stride
    const W* win; i64 wis_b;            // weights with the slices' row pitch, or null (weight = window)
    i32 pitch;                          // row pitch of the slices and weights (elements)
    u32 slice_bytes, wslice_bytes;      // byte extent of one slice and of one weight slice
    void* vol; i32 vs0, vs1, vs2;
    W* wvol; i32 wvs0, wvs1, wvs2;      // null: no output weights
    i32 n;                              // logical size of every axis of the slices, the volume and the target
    i32 nb[3];                          // sub-bricks along z, y, x
    const int4* items;                  // {sub-brick, first slice, end slice, partial slot or -1}
    float* partials;
    const float4* rec;                  // 4 float4 per slice
    i32 n_slices;
    float fbT;                          // window half-width f_b * n, in samples
    float tk;                           // 2 / fbT^2
    float wa[SLAB_NWIN + 1];
    float cut2;                         // cutoff^2 * n^2: a voxel is inside when iz^2 + iy^2 + ix^2 <= cut2
    bool vol_centered;
};

// One record per slice: {R00, R01, R02, -}, {R10, R11, R12, ub}, {R20, R21, R22, uc}, {hw, A, passes, -}. With n =
// the first row (pz = n . v in samples), A = argmax |n_a|, B < C the other axes: the slab |pz| <= fbT runs along A
// around ub * v_B + uc * v_C (ub = -n_B / n_A, uc = -n_C / n_A) with half-width fbT / |n_A|, widened by 2e-3
// samples to cover the rounding of the exact per-voxel test (a few 1e-5 samples). A column holds at most
// floor(2 hw) + 1 candidates: at most 4 when the row is a unit vector (|n_A| >= 1/sqrt(3), fbT <= 1.15), more when
// it is shorter; `passes` = the packing passes of 4 ranks each that the slice needs (1 for rotations).
__global__ __launch_bounds__(256) void slab_records_kernel(const XformArg<float> rot, i32 n_slices, float fbT,
                                                           float4* rec) {
    const i32 s = static_cast<i32>(blockIdx.x) * 256 + static_cast<i32>(threadIdx.x);
    if (s >= n_slices) return;
    float m[9];
    rot_matrix(load_rotation(rot, s), m);      // a quaternion's matrix: its rotation up to rounding
    const float a0 = fabsf(m[0]), a1 = fabsf(m[1]), a2 = fabsf(m[2]);
    i32 ax = 0;
    if (a1 > a0) ax = 1;
    if (a2 > (ax == 0 ? a0 : a1)) ax = 2;
    const float nA = m[ax];
    const float nB = ax == 0 ? m[1] : m[0];
    const float nC = ax == 2 ? m[1] : m[2];
    const float inv = 1.f / nA;
    float4* o = rec + 4 * static_cast<i64>(s);
    o[0] = make_float4(m[0], m[1], m[2], 0.f);
    o[1] = make_float4(m[3], m[4], m[5], -nB * inv);
    o[2] = make_float4(m[6], m[7], m[8], -nC * inv);
    const float hw = fbT * fabsf(inv) + 2e-3f;
    const i32 passes = (static_cast<i32>(fminf(2.f * hw, 15.f)) + 4) >> 2;
    o[3] = make_float4(hw, __int_as_float(ax), __int_as_float(passes), 0.f);
}

// Largest integer m >= 0 with m^2 + r2 <= cut2 (all values exact in single precision), or -1 if none.
__device__ __forceinline__ float sphere_half_extent(float r2, float cut2) {
    const float room = cut2 - r2;
    if (room < 0.f) return -1.f;
    float m = floorf(sqrtf(room));
    if ((m + 1.f) * (m + 1.f) + r2 <= cut2) m += 1.f;
    if (m * m + r2 > cut2) m -= 1.f;
    return m;
}

// Column data of one face orientation for one lane slot: face coordinates (lb, lc), the candidate word of k = 0 and
// the range of k inside the stored volume and the cutoff sphere.
struct SlabColumn {
    float lb, lc, klo, khi;
    u32 word0;
};

// Candidate word: the voxel's index in the sub-brick, (lz * WY + ly) * WX + lx (a 16-bit list entry keeps the
// workgroup's LDS at 13 KB: 12 resident waves per CU).
__device__ __forceinline__ u32 slab_word(i32 lz, i32 ly, i32 lx) {
    return static_cast<u32>((lz * WY + ly) * WX + lx);
}

template<int AX>
__device__ __forceinline__ SlabColumn slab_column(i32 lb, i32 lc, const i32 f0[3], const i32 nin[3], float cut2) {
    // AX 0: k along z, columns (y, x); AX 1: k along y, columns (z, x); AX 2: k along x, columns (z, y).
    SlabColumn c;
    c.lb = static_cast<float>(lb);
    c.lc = static_cast<float>(lc);
    constexpr i32 B = AX == 0 ? 1 : 0;
    constexpr i32 CC = AX == 2 ? 1 : 2;
    const i32 lz = AX == 0 ? 0 : lb, ly = AX == 0 ? lb : (AX == 1 ? 0 : lc), lx = AX == 2 ? 0 : lc;
    c.word0 = slab_word(lz, ly, lx);
    const bool inside = lb < nin[B] && lc < nin[CC];
    const float fb = static_cast<float>(f0[B] + lb), fc = static_cast<float>(f0[CC] + lc);
    const float m = sphere_half_extent(fb * fb + fc * fc, cut2);
    const float fa0 = static_cast<float>(f0[AX]);
    c.klo = fmaxf(-m - fa0, 0.f);
    c.khi = fminf(m - fa0, static_cast<float>(nin[AX] - 1));
    if (!inside || m < 0.f) { c.klo = 1.f; c.khi = 0.f; }
    return c;
}

__device__ __forceinline__ u32 lanes_below(unsigned long long m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<u32>(m), 0u));
}

// Packs the candidates of ranks j0, j0 + 1, ... (at most `jn` ranks) of one face orientation into `cand` (ordered by
// rank, then slot, then lane); returns their number. `u0`: the slab centre along A at face coordinates (0, 0),
// relative to the sub-brick's first sample. Rotations: j0 = 0 and every rank (at most 4 per column, so at most
// 4 x 128 = SLAB_MAXC candidates); thicker slabs: 4 ranks per pass (jn = 4).
template<int AX, int NSLOT>
__device__ __forceinline__ u32 slab_candidates(unsigned short* cand, const SlabColumn (&col)[2], float u0, float ub,
                                               float uc, float hw, i32 j0, i32 jn) {
    constexpr u32 STEP = AX == 0 ? WY * WX : (AX == 1 ? WX : 1);
    float cnt[NSLOT];
    u32 w[NSLOT];
    const float fj0 = static_cast<float>(j0);
#pragma unroll
    for (int s = 0; s < NSLOT; ++s) {
        const float centre = fmaf(ub, col[s].lb, fmaf(uc, col[s].lc, u0));
        const float kl = fmaxf(ceilf(centre - hw), col[s].klo);
        const float kh = fminf(floorf(centre + hw), col[s].khi);
        cnt[s] = (kh - kl) - fj0;                           // candidates from rank j0 on, - 1 (negative: none)
        w[s] = col[s].word0 + static_cast<u32>(static_cast<i32>(fmaxf(kl, 0.f)) + j0) * STEP;
    }
    u32 off = 0;
    for (i32 j = 0; j < 16; ++j) {     // one exit and a constant bound: the compiler unrolls it (faster than a loop)
        bool any = false;
#pragma unroll
        for (int s = 0; s < NSLOT; ++s) {
            const bool has = j < jn && cnt[s] >= static_cast<float>(j);
            const unsigned long long m = __ballot(has);
            if (m == 0) continue;
            any = true;
            if (has) cand[off + lanes_below(m)] = static_cast<unsigned short>(w[s] + static_cast<u32>(j) * STEP);
            off += static_cast<u32>(__popcll(m));
        }
        if (!any) break;
    }
    return off;
}

template<class In, class W>
struct SlabTaps {
    In v00, v01, v10, v11;
    W q00, q01, q10, q11;
    float ty, tx, win;
    u32 lv;
    bool hit, flip, oky0, oky1, okx1;
};

template<class In, class Out, class W, bool SC, bool PAIRED>
__global__ __launch_bounds__(64) void insert_slab_kernel(const SlabParams<In, W> p) {
    constexpr bool CX = vt<In>::cx;
    constexpr i32 NCH = CX ? 3 : 2;                    // accumulator channels: re (, im), weight
    __shared__ float acc[NCH * WVOX];
    __shared__ unsigned short cand[SLAB_MAXC];

    const i32 lane = static_cast<i32>(threadIdx.x);
    int4 item = p.items[blockIdx.x];
    if (item.x < 0) return;                 // padding of an XCD-ordered plan
    item.z = min(item.z, p.n_slices);
    item.y = min(item.y, item.z);
    const i32 brick = item.x;
    const i32 bxi = brick % p.nb[2];
    const i32 rem = brick / p.nb[2];
    const i32 byi = rem % p.nb[1];
    const i32 bzi = rem / p.nb[1];
    const i32 n = p.n, half = n / 2, wr = half + 1;
    const i32 org[3] = {bzi * WZ, byi * WY, bxi * WX};
    const bool vc = p.vol_centered;
    // First frequency of each axis of the sub-brick: frequencies are contiguous inside a sub-brick because n/2 is
    // a multiple of the sub-brick sizes.
    i32 f0[3], nin[3];
    f0[0] = vc ? org[0] - half : (org[0] < half ? org[0] : org[0] - n);
    f0[1] = vc ? org[1] - half : (org[1] < half ? org[1] : org[1] - n);
    f0[2] = org[2];
    nin[0] = WZ;
    nin[1] = WY;
    nin[2] = min(WX, wr - org[2]);

    for (i32 v = lane; v < NCH * WVOX; v += 64) acc[v] = 0.f;

    // Column data of the three face orientations (slot 1 of AX 0 and 1: the columns lane + 64).
    SlabColumn c0[2], c1[2], c2[2];
    // 8 x 8 patches of columns per slot: the 64 lanes of a candidate group project onto a compact patch of the
    // slice, whatever its in-plane rotation (fewer cache lines per tap load than 4 x 16 strips).
    c0[0] = slab_column<0>(lane >> 3, lane & 7, f0, nin, p.cut2);
    c0[1] = slab_column<0>(lane >> 3, (lane & 7) + 8, f0, nin, p.cut2);
    c1[0] = slab_column<1>(lane >> 3, lane & 7, f0, nin, p.cut2);
    c1[1] = slab_column<1>(lane >> 3, (lane & 7) + 8, f0, nin, p.cut2);
    c2[0] = slab_column<2>(lane >> 3, lane & 7, f0, nin, p.cut2);
    c2[1] = c2[0];

    // Culling of a slice against the sub-brick box, in samples: |n . centre| <= fbT + |n| . half + slack.
    const float cz = static_cast<float>(f0[0]) + 0.5f * static_cast<float>(nin[0] - 1);
    const float cy = static_cast<float>(f0[1]) + 0.5f * static_cast<float>(nin[1] - 1);
    const float cx = static_cast<float>(f0[2]) + 0.5f * static_cast<float>(nin[2] - 1);
    const float hz = 0.5f * static_cast<float>(nin[0] - 1), hy = 0.5f * static_cast<float>(nin[1] - 1);
    const float hx = 0.5f * static_cast<float>(nin[2] - 1);
    const float fbT = p.fbT;
    const float f0zf = static_cast<float>(f0[0]), f0yf = static_cast<float>(f0[1]), f0xf = static_cast<float>(f0[2]);

    // Records of the current chunk of 64 slices (lane l: slice chunk0 + l) and of the next one.
    auto load_rec = [&](i32 s, float4 (&r)[4]) {
        if (s < item.z) {
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = p.rec[4 * static_cast<i64>(s) + q];
        } else {
#pragma unroll
            for (int q = 0; q < 4; ++q) r[q] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    float4 cur[4], nxt[4];
    i32 chunk0 = item.y;
    load_rec(chunk0 + lane, cur);
    load_rec(chunk0 + 64 + lane, nxt);
    auto scan = [&]() -> unsigned long long {
        const float d = fabsf((cur[0].x * cz + cur[0].y * cy) + cur[0].z * cx);
        const float bound = fbT + fabsf(cur[0].x) * hz + fabsf(cur[0].y) * hy + fabsf(cur[0].z) * hx;
        return __ballot(chunk0 + lane < item.z && d <= bound + 2e-3f);
    };
    unsigned long long mask = scan();

    const float wa0 = p.wa[0], wa1 = p.wa[1], wa2 = p.wa[2], wa3 = p.wa[3], wa4 = p.wa[4];
    const float wa5 = p.wa[5], wa6 = p.wa[6], wa7 = p.wa[7], wa8 = p.wa[8], wa9 = p.wa[9];
    const float tk = p.tk;
    const i32 pitch = p.pitch;

    // Phase 2 of a group of candidates: interpolation and accumulation (groups of non-hits skip).
    auto finish = [&](const SlabTaps<In, W> (&t)[2]) {
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            if (!t[g].hit) continue;
            const float a0 = t[g].oky0 ? (1.f - t[g].ty) * t[g].win : 0.f;
            const float a1 = t[g].oky1 ? t[g].ty * t[g].win : 0.f;
            const float bx0 = 1.f - t[g].tx, bx1 = t[g].okx1 ? t[g].tx : 0.f;
            const float w00 = a0 * bx0, w01 = a0 * bx1, w10 = a1 * bx0, w11 = a1 * bx1;
            float re, im = 0.f;
            if constexpr (CX) {
                re = fmaf(w11, t[g].v11.x, fmaf(w10, t[g].v10.x, fmaf(w01, t[g].v01.x, w00 * t[g].v00.x)));
                im = fmaf(w11, t[g].v11.y, fmaf(w10, t[g].v10.y, fmaf(w01, t[g].v01.y, w00 * t[g].v00.y)));
                im = t[g].flip ? -im : im;
            } else {
                re = fmaf(w11, t[g].v11, fmaf(w10, t[g].v10, fmaf(w01, t[g].v01, w00 * t[g].v00)));
            }
            const float wt = p.win ? fmaf(w11, t[g].q11, fmaf(w10, t[g].q10, fmaf(w01, t[g].q01, w00 * t[g].q00)))
                                   : t[g].win;
            // A voxel is a candidate at most once per slice, and one wave adds its slices in order: plain
            // read-modify-write.
            acc[t[g].lv] += re;
            if constexpr (CX) acc[WVOX + t[g].lv] += im;
            acc[(NCH - 1) * WVOX + t[g].lv] += wt;
        }
    };
    // Cross-pair pipeline: the taps of the last pair whose sums are not added yet.
    SlabTaps<In, W> pend[2];
    bool have = false;
    i32 pass = 0;          // pass of the current slice (more than one only for rows that are not unit vectors)

    while (true) {
        while (mask == 0) {
            chunk0 += 64;
            if (chunk0 >= item.z) break;
#pragma unroll
            for (int q = 0; q < 4; ++q) cur[q] = nxt[q];
            load_rec(chunk0 + 64 + lane, nxt);
            mask = scan();
        }
        if (mask == 0) break;
        const i32 k = __ffsll(static_cast<long long>(mask)) - 1;
        mask &= mask - 1;
        const i32 s = chunk0 + k;
        // The slice's record, broadcast from the lane that scanned it.
        const float R00 = bcast(cur[0].x, k), R01 = bcast(cur[0].y, k), R02 = bcast(cur[0].z, k);
        const float R10 = bcast(cur[1].x, k), R11 = bcast(cur[1].y, k), R12 = bcast(cur[1].z, k);
        const float R20 = bcast(cur[2].x, k), R21 = bcast(cur[2].y, k), R22 = bcast(cur[2].z, k);
        const float ub = bcast(cur[1].w, k), uc = bcast(cur[2].w, k), hw = bcast(cur[3].x, k);
        const i32 ax = __builtin_amdgcn_readlane(__float_as_int(cur[3].y), k);
        // Rotations pack a slice's candidates in one pass (at most 4 per column). Rows that are not unit vectors
        // thicken the slab: such a slice takes one pass per 4 ranks (a voxel is a candidate at most once per slice,
        // so the passes add distinct voxels); its bit of the mask is set again until the last pass.
        const i32 passes = __builtin_amdgcn_readlane(__float_as_int(cur[3].z), k);
        const i32 j0 = 4 * pass, jn = passes > 1 ? 4 : 16;
        if (passes > 1) {
            if (++pass < passes) mask |= 1ull << k;
            else pass = 0;
        }

        __syncthreads();   // the previous pair's candidate reads are done before the list is rewritten
        u32 total;
        if (ax == 0) {
            const float u0 = fmaf(ub, static_cast<float>(f0[1]), fmaf(uc, static_cast<float>(f0[2]), -f0zf));
            total = slab_candidates<0, 2>(cand, c0, u0, ub, uc, hw, j0, jn);
        } else if (ax == 1) {
            const float u0 = fmaf(ub, static_cast<float>(f0[0]), fmaf(uc, static_cast<float>(f0[2]), -f0yf));
            total = slab_candidates<1, 2>(cand, c1, u0, ub, uc, hw, j0, jn);
        } else {
            const float u0 = fmaf(ub, static_cast<float>(f0[0]), fmaf(uc, static_cast<float>(f0[1]), -f0xf));
            total = slab_candidates<2, 1>(cand, c2, u0, ub, uc, hw, j0, jn);
        }
        __syncthreads();

        const SliceBuf sb(p.in + static_cast<i64>(s) * p.is_b, p.slice_bytes);
        const SliceBuf wb(p.win ? static_cast<const void*>(p.win + static_cast<i64>(s) * p.wis_b)
                                : static_cast<const void*>(p.in), p.win ? p.wslice_bytes : 0u);
        // Phase 1 of candidates [base, base + 128): coordinates, window, and every tap load issued.
        auto issue = [&](u32 base, SlabTaps<In, W> (&t)[2]) {
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                t[g].hit = false;
                if (base + 64u * g >= total) continue;            // uniform
                const u32 pos = base + 64u * g + static_cast<u32>(lane);
                const bool real = pos < total;
                const u32 e = cand[real ? pos : 0u];
                const float iz = f0zf + static_cast<float>(e >> 7);           // WX * WY = 128 voxels per z
                const float iy = f0yf + static_cast<float>((e >> 4) & 7u);
                const float ix = f0xf + static_cast<float>(e & 15u);
                t[g].lv = e;
                // noa's grid2slice in samples (FourierProject.hpp:71-95 with R * (index / n) * n == R * index).
                const float pz = (R00 * iz + R01 * iy) + R02 * ix;
                float py = (R10 * iz + R11 * iy) + R12 * ix;
                float px = (R20 * iz + R21 * iy) + R22 * ix;
                t[g].hit = real && fabsf(pz) <= fbT;
                const float u = pz * pz;
                const float tt = fmaf(u, tk, -1.f);
                float wv = wa9;
                wv = fmaf(wv, tt, wa8); wv = fmaf(wv, tt, wa7); wv = fmaf(wv, tt, wa6); wv = fmaf(wv, tt, wa5);
                wv = fmaf(wv, tt, wa4); wv = fmaf(wv, tt, wa3); wv = fmaf(wv, tt, wa2); wv = fmaf(wv, tt, wa1);
                wv = fmaf(wv, tt, wa0);
                t[g].win = wv;
                // Early flip of LINEAR spectrum interpolation (Interpolation.hpp:459-463).
                t[g].flip = px < 0.f;
                py = t[g].flip ? -py : py;
                px = fabsf(px);
                const float fy = floorf(py), fx = floorf(px);
                t[g].ty = py - fy;
                t[g].tx = px - fx;
                const i32 y0 = static_cast<i32>(fy), x0 = static_cast<i32>(fx);
                // Bounds of the rfft slice (Frequency.hpp:258-288): y in [-n/2, n/2 - 1], x in [0, n/2].
                t[g].oky0 = static_cast<u32>(y0 + half) < static_cast<u32>(n);
                t[g].oky1 = static_cast<u32>(y0 + half) < static_cast<u32>(n - 1);
                t[g].okx1 = x0 < half;
                const i32 row0 = SC ? y0 + half : (y0 & (n - 1));
                const i32 row1 = SC ? row0 + 1 : ((y0 + 1) & (n - 1));
                // Element offsets of the two tap rows (24-bit products: rows and pitch are below 2^24), then byte
                // offsets; the x + 1 taps use the loads' immediate offset.
                const u32 e0 = __umul24(static_cast<u32>(row0), static_cast<u32>(pitch)) + static_cast<u32>(x0);
                const u32 e1 = __umul24(static_cast<u32>(row1), static_cast<u32>(pitch)) + static_cast<u32>(x0);
                // PAIRED lets the compiler merge the two x taps of a row into one load; otherwise the second offset
                // is hidden from it and each tap is its own load.
                const u32 b0 = e0 * static_cast<u32>(sizeof(In)), b1 = e1 * static_cast<u32>(sizeof(In));
                const u32 b0x = PAIRED ? b0 + static_cast<u32>(sizeof(In)) : pin(b0 + static_cast<u32>(sizeof(In)));
                const u32 b1x = PAIRED ? b1 + static_cast<u32>(sizeof(In)) : pin(b1 + static_cast<u32>(sizeof(In)));
                t[g].v00 = sb.template load_bytes<In>(b0);
                t[g].v01 = sb.template load_bytes<In>(b0x);
                t[g].v10 = sb.template load_bytes<In>(b1);
                t[g].v11 = sb.template load_bytes<In>(b1x);
                if (p.win) {
                    const u32 c0 = e0 * 4u, c1 = e1 * 4u;
                    const u32 c0x = PAIRED ? c0 + 4u : pin(c0 + 4u), c1x = PAIRED ? c1 + 4u : pin(c1 + 4u);
                    t[g].q00 = wb.template load_bytes<W>(c0);
                    t[g].q01 = wb.template load_bytes<W>(c0x);
                    t[g].q10 = wb.template load_bytes<W>(c1);
                    t[g].q11 = wb.template load_bytes<W>(c1x);
                }
            }
        };
        SlabTaps<In, W> now[2];
        issue(0u, now);
        // The previous pair's sums are added while this pair's loads are in flight; per voxel, slices still add in
        // order.
        if (have) finish(pend);
        if (total > 128u) {
            finish(now);
            for (u32 base = 128; base < total; base += 128) {
                issue(base, now);
                finish(now);
            }
            have = false;
        } else {
#pragma unroll
            for (int g = 0; g < 2; ++g) pend[g] = now[g];
            have = true;
        }
    }
    if (have) finish(pend);
    __syncthreads();

    if (item.w >= 0) {
        // Split sub-brick: this segment's sums, in the layout of insert_reduce_kernel's partial slots.
        In* pv = reinterpret_cast<In*>(p.partials + static_cast<i64>(item.w) * WVOX * 3);
        W* pw = reinterpret_cast<W*>(p.partials + static_cast<i64>(item.w) * WVOX * 3 + WVOX * 2);
        for (i32 v = lane; v < WVOX; v += 64) {
            if constexpr (CX) pv[v] = make_float2(acc[v], acc[WVOX + v]);
            else pv[v] = acc[v];
            pw[v] = acc[(NCH - 1) * WVOX + v];
        }
        return;
    }
    // Sums of the voxels inside the stored volume and the cutoff are added to the volume (:516-518).
    for (i32 v = lane; v < WVOX; v += 64) {
        const i32 lz = v / (WY * WX), ly = (v / WX) % WY, lx = v % WX;
        if (lx >= nin[2]) continue;
        const float iz = f0zf + static_cast<float>(lz), iy = f0yf + static_cast<float>(ly), ix = f0xf + static_cast<float>(lx);
        if ((iz * iz + iy * iy) + ix * ix > p.cut2) continue;
        const i32 z = org[0] + lz, y = org[1] + ly, x = org[2] + lx;
        In sum;
        if constexpr (CX) sum = make_float2(acc[v], acc[WVOX + v]);
        else sum = acc[v];
        Out* o = static_cast<Out*>(p.vol) + (z * p.vs0 + y * p.vs1 + x * p.vs2);
        *o = vadd(*o, vcast<Out>(sum));
        if (p.wvol) {
            W* w = p.wvol + (z * p.wvs0 + y * p.wvs1 + x * p.wvs2);
            *w = *w + acc[(NCH - 1) * WVOX + v];
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Pose-adaptive work plan (item order 3): how many slices pass each active sub-brick, in bins of the slice range
// (count_bin_width slices each, at most CNT_BINS), counted with the one-wave kernel's culling test
// (|n . centre| <= f_b + |n| . half-extent, normalised frequencies: a superset of the slices with hits). The host
// cuts each sub-brick's range at bin boundaries into segments of about the target number of passing slices, so that
// a sub-brick crossed by far more slices than the uniform-orientation model expects (tilt series, preferred
// orientations) is split instead of becoming one long item.
// ---------------------------------------------------------------------------------------------------------
constexpr i32 CNT_BINS = 64;

template<class C>
struct CountParams {
    XformArg<C> rot;
    const i32* bricks;                  // active sub-bricks
    i32* counts;                        // CNT_BINS per active sub-brick
    i32 n_slices, n_bricks;
    i32 bin_w;                          // slices per bin: a multiple of 64, so that 64 consecutive slices share a bin
    i32 range_w;                        // slices per workgroup range (blockIdx.y), a multiple of 64
    i32 grid[3], nb[3], bd[3];
    C target[3], fb;
    bool vol_centered;
};

// Slices per bin of the pass counts: 64 * ceil(n / (64 * CNT_BINS)).
inline i32 count_bin_width(i32 n_slices) {
    return 64 * std::max<i32>(1, (n_slices + 64 * CNT_BINS - 1) / (64 * CNT_BINS));
}

constexpr i32 CNT_BRICKS = 16;      // sub-bricks per wave of the count kernel (each rotation read once for all)

template<class C>
__global__ __launch_bounds__(64) void insert_count_kernel(const CountParams<C> p) {
    __shared__ i32 bins[CNT_BRICKS][CNT_BINS];
    __shared__ C box[CNT_BRICKS][6];                   // centre and half-extent of each sub-brick
    const i32 lane = static_cast<i32>(threadIdx.x);
    const i32 j0 = static_cast<i32>(blockIdx.x) * CNT_BRICKS;
    const i32 nj = min(CNT_BRICKS, p.n_bricks - j0);
    for (i32 i = lane; i < CNT_BRICKS * CNT_BINS; i += 64) bins[i / CNT_BINS][i % CNT_BINS] = 0;
    if (lane < nj) {
        const i32 brick = p.bricks[j0 + lane];
        const i32 bx = brick % p.nb[2], r = brick / p.nb[2], by = r % p.nb[1], bz = r / p.nb[1];
        const i32 org[3] = {bz * p.bd[0], by * p.bd[1], bx * p.bd[2]};
        const i32 ext[3] = {p.grid[0], p.grid[1], p.grid[2] / 2 + 1};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            i32 lo = 0x7fffffff, hi = -0x7fffffff;
            for (i32 k = 0; k < p.bd[a] && org[a] + k < ext[a]; ++k) {
                const i32 f = axis_freq(a == 2, org[a] + k, p.vol_centered, p.grid[a]);
                lo = min(lo, f);
                hi = max(hi, f);
            }
            box[lane][a] = (static_cast<C>(lo) + static_cast<C>(hi)) * C(0.5) / p.target[a];
            box[lane][3 + a] = (static_cast<C>(hi) - static_cast<C>(lo)) * C(0.5) / p.target[a];
        }
    }
    __syncthreads();
    // This workgroup's range of slices, 64 consecutive slices per step, all in bin s0 / bin_w (bin_w and range_w are
    // multiples of 64).
    const i32 r0 = static_cast<i32>(blockIdx.y) * p.range_w, r1 = min(p.n_slices, r0 + p.range_w);
    for (i32 s0 = r0; s0 < r1; s0 += 64) {
        const i32 s = s0 + lane;
        const bool live = s < p.n_slices;
        C m[9] = {};
        if (live) rot_matrix(load_rotation(p.rot, s), m);
        const i32 bin = s0 / p.bin_w;
        for (i32 j = 0; j < nj; ++j) {
            const C d = fabs(m[0] * box[j][0] + m[1] * box[j][1] + m[2] * box[j][2]);
            const C bound = p.fb + fabs(m[0]) * box[j][3] + fabs(m[1]) * box[j][4] + fabs(m[2]) * box[j][5];
            const unsigned long long pass = __ballot(live && d <= bound + C(1e-5));
            if (lane == 0 && pass != 0) bins[j][bin] += __popcll(pass);
        }
    }
    __syncthreads();
    for (i32 i = lane; i < nj * CNT_BINS; i += 64) {
        const i32 c = bins[i / CNT_BINS][i % CNT_BINS];
        if (c != 0) atomicAdd(p.counts + (static_cast<i64>(j0 + i / CNT_BINS) * CNT_BINS + i % CNT_BINS), c);
    }
}

__global__ __launch_bounds__(256) void zero_i32_kernel(i32* __restrict__ a, i64 n) {
    const i64 i = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) a[i] = 0;
}

// General path: one lane per voxel, every slice in order (noa's operator as is).
template<class C, class In, class Out, class W, int INTERP, bool GENERAL>
__global__ __launch_bounds__(256) void insert_general_kernel(const InsertParams<C, In, W> p) {
    using RIn = typename vt<In>::real;
    const i32 nxy = p.grid[1] * p.wr;
    const i64 t = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= static_cast<i64>(p.grid[0]) * nxy) return;
    const i32 z = static_cast<i32>(t / nxy);
    const i32 r = static_cast<i32>(t - static_cast<i64>(z) * nxy);
    const i32 y = r / p.wr, x = r - (r / p.wr) * p.wr;

    vec3<C> f;
    f.z = static_cast<C>(index2frequency(p.vol_centered, z, p.grid[0])) / p.target[0];
    f.y = static_cast<C>(index2frequency(p.vol_centered, y, p.grid[1])) / p.target[1];
    f.x = static_cast<C>(x) / p.target[2];
    if (dot3(f) > p.cutoff_sqd) return;

    In value = vzero<In>();
    W weights{};
    for (i32 i = p.slice_begin; i < p.slice_end; ++i) {
        const Rot<C> rot = load_rotation(p.rot, i);
        Scale<C> sc;
        sc.identity = true;
        if constexpr (GENERAL) sc = load_scaling(p.scale, i);
        vec2<C> f2;
        const C fz = grid2slice<GENERAL>(f, sc, rot, p.ews[0], p.ews[1], f2);
        if (fabs(fz) <= p.fftfreq_blackman) {
            const C window = windowed_sinc<C>(fz, p.fftfreq_sinc, p.fftfreq_blackman);
            const C py = f2.y * p.slice_f[0], px = f2.x * p.slice_f[1];
            const In* base = p.in + static_cast<i64>(i) * p.is_b;
            const In v = interp2<INTERP, In>(base, p.is_y, p.is_x, p.slice_shape[0], p.slice_shape[1],
                                             p.slices_centered, py, px, p.in_is_value, p.in_value, p.interp_mode);
            value = vadd(value, vmul(v, static_cast<RIn>(window)));
            if (p.has_wvol) {
                if (p.win_present) {
                    const W* wbase = p.win + static_cast<i64>(i) * p.wis_b;
                    const W wv = interp2<INTERP, W>(wbase, p.wis_y, p.wis_x, p.slice_shape[0], p.slice_shape[1],
                                                    p.slices_centered, py, px, p.win_is_value, p.win_value,
                                                    p.interp_mode);
                    weights = weights + wv * static_cast<W>(window);
                } else {
                    weights = weights + static_cast<W>(window);
                }
            }
        }
    }
    Out* vol = static_cast<Out*>(p.vol) + (z * p.vs0 + y * p.vs1 + x * p.vs2);
    *vol = vadd(*vol, vcast<Out>(value));
    if (p.has_wvol) {
        W* wv = p.wvol + (z * p.wvs0 + y * p.wvs1 + x * p.wvs2);
        *wv = *wv + weights;
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

i32 host_freq(bool centered, i32 i, i32 n) { return centered ? i - n / 2 : (i < (n + 1) / 2 ? i : i - n); }

// Polynomial form of windowed_sinc(f, fs, fb) (FourierProject.hpp:99-110) on f in [-fb, fb], in t = 2 (f/fb)^2 - 1
// (the window is even): Chebyshev interpolation of degree NWIN at the Chebyshev nodes, converted to monomials in
// t and rounded to single precision. The fit is accepted when the kernel's own single-precision evaluation
// (window_poly, reproduced here with the same roundings) stays within 4e-7 of the double-precision window on a
// dense grid; otherwise the kernel evaluates noa's expression. Kept per (fs, fb) for the process lifetime.
struct WindowFit {
    float a[NWIN + 1];
    bool ok;
};
std::mutex g_window_mutex;
std::map<std::pair<float, float>, WindowFit> g_windows;

WindowFit fit_window(float fs, float fb) {
    std::lock_guard<std::mutex> lock(g_window_mutex);
    const auto key = std::make_pair(fs, fb);
    auto it = g_windows.find(key);
    if (it != g_windows.end()) return it->second;
    const double PI = 3.141592653589793238462643383279502884;
    auto w = [&](double f) {
        const double x = PI * f / fs;
        const double s = x == 0 ? 1.0 : std::sin(x) / x;
        const double b = PI * f / fb;
        return s * (0.42 + 0.5 * std::cos(b) + 0.08 * std::cos(2 * b));
    };
    constexpr i32 n = NWIN, m = NWIN + 1;
    double c[NWIN + 1] = {};
    for (i32 k = 0; k <= n; ++k) {
        double sum = 0;
        for (i32 j = 0; j < m; ++j) {
            const double th = PI * (j + 0.5) / m;
            const double u = 0.5 * (std::cos(th) + 1);
            sum += w(static_cast<double>(fb) * std::sqrt(u)) * std::cos(k * th);
        }
        c[k] = 2.0 * sum / m;
    }
    c[0] *= 0.5;
    // T_k(t) in monomials: T_{k+1} = 2 t T_k - T_{k-1}.
    double tk[NWIN + 1][NWIN + 1] = {};
    tk[0][0] = 1;
    tk[1][1] = 1;
    for (i32 k = 2; k <= n; ++k)
        for (i32 j = 0; j <= k; ++j) tk[k][j] = (j > 0 ? 2 * tk[k - 1][j - 1] : 0.0) - tk[k - 2][j];
    WindowFit fit{};
    for (i32 j = 0; j <= n; ++j) {
        double a = 0;
        for (i32 k = j; k <= n; ++k) a += c[k] * tk[k][j];
        fit.a[j] = static_cast<float>(a);
    }
    const float inv_fb = 1.f / fb;
    double err = 0;
    for (i32 j = -4000; j <= 4000; ++j) {
        const float f = fb * (static_cast<float>(j) / 4000.f);
        const float q = f * inv_fb;
        const float t = std::fmaf(2.f * q, q, -1.f);
        float r = fit.a[n];
        for (i32 k = n - 1; k >= 0; --k) r = std::fmaf(r, t, fit.a[k]);
        err = std::max(err, std::abs(static_cast<double>(r) - w(static_cast<double>(f))));
    }
    fit.ok = err <= 4e-7;
    g_windows[key] = fit;
    return fit;
}

// Work plan of the culled kernels for one geometry: the active bricks (box within the cutoff sphere), each cut into
// slice segments so that the expected number of passing slices per item is at most `list_target`. The expected
// count assumes orientations uniform on the sphere: a brick at distance r meets a fraction min(1, t / r) of the
// slabs, t = f_b + mean projected half-extent. Items are sorted by expected work, largest first. Kept per (device,
// geometry, brick shape, slice count, window) for the process lifetime.
struct Plan {
    int4* items{};                      // device copies, uploaded on first use by the allocating entry point
    int4* reduce{};
    i32 n_items{}, n_reduce{}, n_partials{};
    std::vector<int4> h_items, h_reduce;   // host copies (the allocation-free entry point copies them to its scratch)
    std::vector<float> h_work;             // expected passing slices of each item
    std::vector<i32> h_bricks;             // active sub-bricks (pose-adaptive plans count their passing slices)
    std::vector<i32> h_nseg;               // segments of each active sub-brick in this plan
    i32* bricks{};                         // device copy of h_bricks, uploaded on first use
    bool owned{};                          // device arrays freed with the plan (pose-adaptive plans, one call)
    double target{};                       // expected passing slices per item the plan was cut for
};

// The default work-item size adapts to the call: at most the kernel's value, and small enough that the call's
// expected passing (sub-brick, slice) pairs fill about this many items (a few per resident wave of the GPU), so that
// calls of a few hundred or thousand slices keep every CU busy instead of a few long items.
constexpr double PLAN_MIN_ITEMS = 4096;
using PlanKey = std::tuple<int, i32, i32, i32, i64, i64, i64, double, bool, i32, double, i32, i32, i32, double, int, i32>;
std::mutex g_plan_mutex;
std::map<PlanKey, Plan> g_plans;

// Device copies of a plan's items and reductions (once per plan; called with g_plan_mutex held).
int upload_plan(Plan& pl, std::string& err) {
    if (!pl.h_items.empty() && !pl.items) {
        if (hipMalloc(&pl.items, pl.h_items.size() * sizeof(int4)) != hipSuccess ||
            hipMemcpy(pl.items, pl.h_items.data(), pl.h_items.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) {
            err = "insert: could not upload the work plan";
            return NOA_HIP_ERROR_RUNTIME;
        }
    }
    if (!pl.h_reduce.empty() && !pl.reduce) {
        if (hipMalloc(&pl.reduce, pl.h_reduce.size() * sizeof(int4)) != hipSuccess ||
            hipMemcpy(pl.reduce, pl.h_reduce.data(), pl.h_reduce.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess) {
            err = "insert: could not upload the work plan";
            return NOA_HIP_ERROR_RUNTIME;
        }
    }
    return NOA_HIP_OK;
}

// The plan of one geometry, built on first use and kept (map nodes are stable, so `plan` stays valid). `upload`:
// also make the device copies (the allocation-free entry point only reads the host copies).
// order: 0 = items by expected work, largest first; 1 = by first slice, then by expected work (items running at the
// same time read neighbouring slices: their footprints are shared in the L2); 2 = XCD-aware spatial order (below).
int make_plan(int device, const i32 grid[3], const i64 target[3], double cutoff, bool centered, i32 n_slices,
              double fftfreq_blackman, const i32 bd[3], double list_target, bool upload, const Plan*& plan,
              std::string& err, int order = 0, i32 max_span = 0) {
    const PlanKey key{device, grid[0], grid[1], grid[2], target[0], target[1], target[2], cutoff, centered, n_slices,
                      fftfreq_blackman, bd[0], bd[1], bd[2], list_target, order, max_span};
    std::lock_guard<std::mutex> lock(g_plan_mutex);
    auto it = g_plans.find(key);
    if (it != g_plans.end()) {
        if (upload) {
            const int st = upload_plan(it->second, err);
            if (st != NOA_HIP_OK) return st;
        }
        plan = &it->second;
        return NOA_HIP_OK;
    }
    const i32 ext[3] = {grid[0], grid[1], grid[2] / 2 + 1};
    const i32 nb[3] = {(ext[0] + bd[0] - 1) / bd[0], (ext[1] + bd[1] - 1) / bd[1], (ext[2] + bd[2] - 1) / bd[2]};
    const double c2 = std::max(cutoff, 0.0) * std::max(cutoff, 0.0);
    struct Entry { double work; int4 item; };
    struct Active { i32 brick; double expected; };
    std::vector<Active> active;
    std::vector<Entry> entries;
    std::vector<int4> reduce;
    i32 n_partials = 0;
    for (i32 bz = 0; bz < nb[0]; ++bz)
        for (i32 by = 0; by < nb[1]; ++by)
            for (i32 bx = 0; bx < nb[2]; ++bx) {
                const i32 org[3] = {bz * bd[0], by * bd[1], bx * bd[2]};
                double lo[3], hi[3];
                for (int a = 0; a < 3; ++a) {
                    lo[a] = 1e300; hi[a] = -1e300;
                    for (i32 k = 0; k < bd[a] && org[a] + k < ext[a]; ++k) {
                        const i32 f = a == 2 ? org[a] + k : host_freq(centered, org[a] + k, grid[a]);
                        const double fn = static_cast<double>(f) / static_cast<double>(target[a]);
                        lo[a] = std::min(lo[a], fn);
                        hi[a] = std::max(hi[a], fn);
                    }
                }
                double closest = 0, r2 = 0, t = fftfreq_blackman;
                for (int a = 0; a < 3; ++a) {
                    const double d = lo[a] > 0 ? lo[a] : (hi[a] < 0 ? -hi[a] : 0.0);
                    closest += d * d;
                    const double c = 0.5 * (lo[a] + hi[a]);
                    r2 += c * c;
                    t += 0.25 * (hi[a] - lo[a]);   // E|n_a| = 1/2 times the half-extent
                }
                if (closest > c2 * (1 + 1e-6) + 1e-12) continue;
                const double density = std::min(1.0, t / std::max(std::sqrt(r2), 1e-12));
                active.push_back({(bz * nb[1] + by) * nb[2] + bx, density * n_slices});
            }
    // A negative list_target is the default: at most -list_target slices per item, fewer for small calls.
    double per_item = list_target;
    if (list_target < 0) {
        double total = 0;
        for (const Active& a : active) total += a.expected;
        per_item = std::min(-list_target, std::max(16.0, total / PLAN_MIN_ITEMS));
    }
    for (const Active& a : active) {
                const i32 brick = a.brick;
                const double expected = a.expected;
                // At most max_span slices per segment (0: no bound): a sub-brick crossed by many more slices than
                // the model expects (orientation sets far from uniform) then still gets several items.
                const i32 span_segs = max_span > 0 ? (n_slices + max_span - 1) / max_span : 1;
                const i32 n_seg = std::max<i32>(span_segs, static_cast<i32>(std::ceil(expected / per_item)));
                const i32 len = (n_slices + n_seg - 1) / n_seg;
                const i32 first_slot = n_seg > 1 ? n_partials : -1;
                i32 made = 0;
                for (i32 sgm = 0; sgm < n_seg; ++sgm) {
                    const i32 b = sgm * len, e = std::min(n_slices, b + len);
                    if (b >= e) break;
                    entries.push_back({expected / n_seg, make_int4(brick, b, e, n_seg > 1 ? n_partials + made : -1)});
                    ++made;
                }
                if (n_seg > 1) {
                    reduce.push_back(make_int4(brick, first_slot, made, 0));
                    n_partials += made;
                }
            }
    if (order == 1) {
        std::stable_sort(entries.begin(), entries.end(), [](const Entry& a, const Entry& b) {
            return a.item.y != b.item.y ? a.item.y < b.item.y : a.work > b.work;
        });
    } else if (order == 2) {
        // XCD-aware: workgroup i runs on XCD i % 8. The sub-bricks are put in Morton order and cut into 8 contiguous
        // runs of equal expected work; item i of the array comes from run i % 8, in order, so each XCD walks one
        // compact region and the items resident on it at a time are neighbours reading the same slices (their
        // footprints shared in that XCD's L2). Runs shorter than the longest are padded with no-op items (x = -1).
        auto morton = [&](i32 brick) {
            const i32 bx = brick % nb[2], r = brick / nb[2], by = r % nb[1], bz = r / nb[1];
            unsigned long long c = 0;
            for (int k = 0; k < 11; ++k)
                c |= (static_cast<unsigned long long>((bz >> k) & 1) << (3 * k + 2)) |
                     (static_cast<unsigned long long>((by >> k) & 1) << (3 * k + 1)) |
                     (static_cast<unsigned long long>((bx >> k) & 1) << (3 * k));
            return c;
        };
        std::stable_sort(entries.begin(), entries.end(), [&](const Entry& a, const Entry& b) {
            const unsigned long long ma = morton(a.item.x), mb = morton(b.item.x);
            return ma != mb ? ma < mb : a.item.y < b.item.y;
        });
        double total = 0;
        for (const Entry& e : entries) total += e.work;
        std::vector<Entry> runs[8];
        double cum = 0;
        for (const Entry& e : entries) {
            const int r = total > 0 ? std::min(7, static_cast<int>(8 * (cum + 0.5 * e.work) / total)) : 0;
            runs[r].push_back(e);
            cum += e.work;
        }
        size_t longest = 0;
        for (const auto& r : runs) longest = std::max(longest, r.size());
        std::vector<Entry> out;
        out.reserve(8 * longest);
        for (size_t j = 0; j < longest; ++j)
            for (const auto& r : runs) out.push_back(j < r.size() ? r[j] : Entry{0.0, make_int4(-1, 0, 0, -1)});
        entries.swap(out);
    } else {
        std::stable_sort(entries.begin(), entries.end(), [](const Entry& a, const Entry& b) { return a.work > b.work; });
    }
    std::vector<int4> items(entries.size());
    std::vector<float> work(entries.size());
    for (size_t k = 0; k < entries.size(); ++k) {
        items[k] = entries[k].item;
        work[k] = static_cast<float>(entries[k].work);
    }
    std::map<i32, i32> segments_of;
    for (const int4& it : items) ++segments_of[it.x];
    Plan& pl = g_plans[key];
    pl.target = per_item;
    for (const auto& kv : segments_of) {
        pl.h_bricks.push_back(kv.first);
        pl.h_nseg.push_back(kv.second);
    }
    pl.n_items = static_cast<i32>(items.size());
    pl.n_reduce = static_cast<i32>(reduce.size());
    pl.n_partials = n_partials;
    pl.h_items = std::move(items);
    pl.h_work = std::move(work);
    pl.h_reduce = std::move(reduce);
    plan = &pl;
    return upload ? upload_plan(pl, err) : NOA_HIP_OK;
}

// The pose-adaptive plan of one call: for each active sub-brick (counts: CNT_BINS passing-slice counts over the slice
// range), segments cut at bin boundaries with about `list_target` passing slices each; sub-bricks crossed by no slice
// get no item. Items sorted by passing slices, largest first; partial slots and reductions as make_plan's.
void make_pose_plan(const std::vector<i32>& bricks, const std::vector<i32>& counts, i32 n_slices, double list_target,
                    Plan& pl) {
    struct Entry { double work; int4 item; };
    std::vector<Entry> entries;
    entries.reserve(bricks.size() * 2);
    std::vector<int4> reduce;
    i32 n_partials = 0;
    const i32 bin_w = count_bin_width(n_slices);
    auto bin_start = [&](i32 b) { return std::min<i32>(n_slices, b * bin_w); };
    i32 seg_lo[CNT_BINS], seg_hi[CNT_BINS];
    i64 seg_work[CNT_BINS];
    for (size_t j = 0; j < bricks.size(); ++j) {
        const i32* c = counts.data() + j * CNT_BINS;
        i64 total = 0;
        for (i32 b = 0; b < CNT_BINS; ++b) total += c[b];
        if (total == 0) continue;
        const i64 n_seg = std::max<i64>(1, static_cast<i64>(std::ceil(static_cast<double>(total) / list_target)));
        // Segments cut where the running count passes k / n_seg of the total, each trimmed to its first and last
        // non-empty bins.
        i32 nk = 0;
        i64 acc = 0, prev = 0;
        i32 lo = -1, hi = -1;
        for (i32 b = 0; b < CNT_BINS; ++b) {
            if (c[b] > 0) {
                if (lo < 0) lo = b;
                hi = b;
            }
            acc += c[b];
            const bool cut = b == CNT_BINS - 1 || (nk < n_seg - 1 && acc * n_seg >= total * (nk + 1));
            if (cut && lo >= 0) {
                seg_lo[nk] = lo;
                seg_hi[nk] = hi + 1;
                seg_work[nk] = acc - prev;
                ++nk;
                prev = acc;
                lo = hi = -1;
            }
        }
        const bool split = nk > 1;
        for (i32 k = 0; k < nk; ++k)
            entries.push_back({static_cast<double>(seg_work[k]),
                               make_int4(bricks[j], bin_start(seg_lo[k]), bin_start(seg_hi[k]),
                                         split ? n_partials + k : -1)});
        if (split) {
            reduce.push_back(make_int4(bricks[j], n_partials, nk, 0));
            n_partials += nk;
        }
    }
    std::stable_sort(entries.begin(), entries.end(), [](const Entry& a, const Entry& b) { return a.work > b.work; });
    pl.h_items.resize(entries.size());
    pl.h_work.resize(entries.size());
    for (size_t k = 0; k < entries.size(); ++k) {
        pl.h_items[k] = entries[k].item;
        pl.h_work[k] = static_cast<float>(entries[k].work);
    }
    pl.h_reduce = std::move(reduce);
    pl.n_items = static_cast<i32>(pl.h_items.size());
    pl.n_reduce = static_cast<i32>(pl.h_reduce.size());
    pl.n_partials = n_partials;
}

// Culled kernel of the insertion: the slab kernel where the configuration allows it, else the one-wave kernel
// (NOA_HIP_INSERT_KERNEL=wave forces the one-wave kernel), and the expected passing slices per item.
struct KernelChoice {
    bool force_wave;
    bool force_general;   // NOA_HIP_INSERT_KERNEL=general: noa's per-voxel gather over every slice
    bool paired;          // slab kernel: one 16-byte (8-byte) load for the two x taps of a slice (weight) row
    int order;            // order of the work items (make_plan)
    double list_target;   // expected passing slices per work item; 0: the kernel's default
    i32 max_span;         // most slices per work item (0: no bound; not for the allocation-free entry point or the
                          // pose-adaptive order, whose segments come from the counts)
};
KernelChoice kernel_choice(const ProjectDesc& d) {
    KernelChoice k{};
    i32 v[4];
    const int remap = (d.src_centered ? 4 : 0) + (d.dst_centered ? 8 : 0);
    tuning_lookup(NOA_HIP_OP_INSERT, d.out_shape[1], remap, d.interp, d.output.dtype, v);
    // The slab kernel where it applies (5.5 against 6.0 ms at 256^3 with 10,000 slices), unless the table asks for
    // the one-wave kernel (1); its two x taps of a row as one load unless the table says 2.
    k.force_wave = v[0] == 1;
    k.list_target = v[1] > 0 ? std::max(16.0, static_cast<double>(v[1])) : 0.0;   // 0: the kernel's default
    k.paired = v[2] != 2;
    k.order = v[3];
    if (const char* o = std::getenv("NOA_HIP_INSERT_ORDER")) k.order = std::atoi(o);
    if (const char* e = std::getenv("NOA_HIP_INSERT_KERNEL")) {
        k.force_wave = std::string(e) != "slab";
        k.force_general = std::string(e) == "general";
    }
    if (const char* pr = std::getenv("NOA_HIP_INSERT_PAIRED")) k.paired = std::string(pr) != "0";
    if (const char* l = std::getenv("NOA_HIP_INSERT_LIST")) k.list_target = std::max(16.0, std::atof(l));
    // At most 1,024 slices per work item: a sub-brick crossed by many more slices than the uniform model expects
    // (a tilt series: 13.4 -> 6.1 ms at 256^3 with 10,000 slices) still gets several items, for uniform poses within
    // the noise (more partial sums, 591 against 140 MiB at 10,000 slices).
    k.max_span = 1024;
    if (const char* sp = std::getenv("NOA_HIP_INSERT_SPAN")) k.max_span = std::max(0, std::atoi(sp));
    return k;
}

// Window of the slab kernel: windowed_sinc(pz / n, fs, fb) (FourierProject.hpp:99-110) as a polynomial of degree
// SLAB_NWIN in t = 2 (pz / fbT)^2 - 1, pz in samples (fbT = fb * n): Chebyshev interpolation at the Chebyshev nodes,
// converted to monomials in t and rounded to single precision. Accepted when the kernel's own evaluation (u = pz^2,
// t = fma(u, 2 / fbT^2, -1), Horner with fma, reproduced here) stays within 4e-7 of the double-precision window.
struct SlabWindow {
    float a[SLAB_NWIN + 1];
    float fbT, tk;
    bool ok;
};
std::map<std::tuple<float, float, i32>, SlabWindow> g_slab_windows;

SlabWindow fit_slab_window(float fs, float fb, i32 n) {
    std::lock_guard<std::mutex> lock(g_window_mutex);
    const auto key = std::make_tuple(fs, fb, n);
    auto it = g_slab_windows.find(key);
    if (it != g_slab_windows.end()) return it->second;
    const double PI = 3.141592653589793238462643383279502884;
    auto w = [&](double f) {
        const double x = PI * f / fs;
        const double s = x == 0 ? 1.0 : std::sin(x) / x;
        const double b = PI * f / fb;
        return s * (0.42 + 0.5 * std::cos(b) + 0.08 * std::cos(2 * b));
    };
    SlabWindow sw{};
    sw.fbT = fb * static_cast<float>(n);
    sw.tk = 2.f / (sw.fbT * sw.fbT);
    constexpr i32 deg = SLAB_NWIN, m = SLAB_NWIN + 1;
    double c[SLAB_NWIN + 1] = {};
    for (i32 k = 0; k <= deg; ++k) {
        double sum = 0;
        for (i32 j = 0; j < m; ++j) {
            const double th = PI * (j + 0.5) / m;
            const double u = 0.5 * (std::cos(th) + 1);                 // (pz / fbT)^2
            sum += w(static_cast<double>(fb) * std::sqrt(u)) * std::cos(k * th);
        }
        c[k] = 2.0 * sum / m;
    }
    c[0] *= 0.5;
    double tkm[SLAB_NWIN + 1][SLAB_NWIN + 1] = {};
    tkm[0][0] = 1;
    tkm[1][1] = 1;
    for (i32 k = 2; k <= deg; ++k)
        for (i32 j = 0; j <= k; ++j) tkm[k][j] = (j > 0 ? 2 * tkm[k - 1][j - 1] : 0.0) - tkm[k - 2][j];
    for (i32 j = 0; j <= deg; ++j) {
        double a = 0;
        for (i32 k = j; k <= deg; ++k) a += c[k] * tkm[k][j];
        sw.a[j] = static_cast<float>(a);
    }
    double err = 0;
    for (i32 j = -4000; j <= 4000; ++j) {
        const float pz = sw.fbT * (static_cast<float>(j) / 4000.f);
        const float u = pz * pz;
        const float t = std::fmaf(u, sw.tk, -1.f);
        float r = sw.a[deg];
        for (i32 k = deg - 1; k >= 0; --k) r = std::fmaf(r, t, sw.a[k]);
        err = std::max(err, std::abs(static_cast<double>(r) - w(static_cast<double>(pz) / n)));
    }
    sw.ok = err <= 4e-7;
    g_slab_windows[key] = sw;
    return sw;
}

// Device scratch of the allocation-free entry point: the work plan, the slice records and the partial sums of one
// call are carved from it (256-byte aligned pieces) instead of being allocated.
struct Scratch {
    char* base;
    size_t bytes;
};

// The kernel parameters of one call (everything but the work plan and the partial sums); returns whether the slice
// extents fit the 32-bit buffer loads of the culled kernels.
template<class C, class In, class W>
bool fill_params(const ProjectDesc& d, InsertParams<C, In, W>& p) {
    p = InsertParams<C, In, W>{};
    p.in = static_cast<const In*>(d.input.data);
    p.is_b = d.input.s0; p.is_y = (i32)d.input.s1; p.is_x = (i32)d.input.s2;
    if (d.in_shape[0] == 1) p.is_b = 0;   // batch broadcast of the interpolator input (xform/Utils.hpp:25-28)
    p.in_is_value = d.input.is_value; p.in_value = value_of<In>(d.input);
    p.win_present = d.input_weight.present();
    p.win = static_cast<const W*>(d.input_weight.data);
    p.wis_b = d.in_shape[0] == 1 ? 0 : d.input_weight.s0; p.wis_y = (i32)d.input_weight.s1; p.wis_x = (i32)d.input_weight.s2;
    p.win_is_value = d.input_weight.is_value; p.win_value = static_cast<W>(d.input_weight.value[0]);
    p.vol = d.output.data; p.vs0 = (i32)d.output.s0; p.vs1 = (i32)d.output.s1; p.vs2 = (i32)d.output.s2;
    p.has_wvol = d.output_weight.data != nullptr;
    p.wvol = static_cast<W*>(d.output_weight.data);
    p.wvs0 = (i32)d.output_weight.s0; p.wvs1 = (i32)d.output_weight.s1; p.wvs2 = (i32)d.output_weight.s2;
    p.pow2 = true;
    for (int i = 0; i < 3; ++i) {
        p.grid[i] = static_cast<i32>(d.out_shape[i + 1]);
        p.target[i] = static_cast<C>(d.target[i]);
        const i64 t = d.target[i];
        p.pow2 = p.pow2 && t > 0 && (t & (t - 1)) == 0;
        p.inv_target[i] = C(1) / p.target[i];
    }
    p.wr = p.grid[2] / 2 + 1;
    p.bd[0] = WZ;
    p.bd[1] = WY;
    p.bd[2] = WX;
    p.apitch = WX;
    p.nb[0] = (p.grid[0] + p.bd[0] - 1) / p.bd[0];
    p.nb[1] = (p.grid[1] + p.bd[1] - 1) / p.bd[1];
    p.nb[2] = (p.wr + p.bd[2] - 1) / p.bd[2];
    p.slice_f[0] = static_cast<C>(d.in_shape[2]);
    p.slice_f[1] = static_cast<C>(d.in_shape[3]);
    p.slice_shape[0] = static_cast<i32>(d.in_shape[2]);
    p.slice_shape[1] = static_cast<i32>(d.in_shape[3]);
    p.n_slices = static_cast<i32>(d.in_shape[0]);
    const C cutoff = std::max(static_cast<C>(d.cutoff), C(0));
    p.cutoff_sqd = cutoff * cutoff;
    // Clamp the windowed sinc to at least one pixel (:471-474).
    const C max_output_size = static_cast<C>(std::min({d.target[0], d.target[1], d.target[2]}));
    p.fftfreq_sinc = std::max(static_cast<C>(d.fftfreq_sinc), 1 / max_output_size);
    p.fftfreq_blackman = std::max(static_cast<C>(d.fftfreq_blackman), 1 / max_output_size);
    p.general = d.general;
    if (d.general && (d.ews_radius[0] != 0 || d.ews_radius[1] != 0)) {
        p.ews[0] = 1 / (2 * static_cast<C>(d.ews_radius[0]));
        p.ews[1] = 1 / (2 * static_cast<C>(d.ews_radius[1]));
    }
    p.rot = to_arg<C>(d.rot);
    p.scale = to_arg<C>(d.scale);
    p.slices_centered = d.src_centered;
    p.vol_centered = d.dst_centered;
    {
        const WindowFit fit = fit_window(static_cast<float>(p.fftfreq_sinc), static_cast<float>(p.fftfreq_blackman));
        for (int k = 0; k <= NWIN; ++k) p.wa[k] = fit.a[k];
        p.wpoly = fit.ok;
        p.inv_fb = 1.f / static_cast<float>(p.fftfreq_blackman);
    }
    p.slice_begin = 0;
    p.slice_end = p.n_slices;

    // Byte extents of one slice and of its weights, for the buffer loads of the one-wave kernel.
    auto extent = [&](i64 sy, i64 sx, size_t elem) -> u64 {
        if (sy < 0 || sx < 0) return ~u64(0);
        return (static_cast<u64>(p.slice_shape[0] - 1) * sy + static_cast<u64>(p.slice_shape[1] / 2) * sx + 1) * elem;
    };
    const u64 sb = extent(p.is_y, p.is_x, sizeof(In));
    const u64 wb = p.win_present ? extent(p.wis_y, p.wis_x, sizeof(W)) : 0;
    const bool buffers_ok = sb < 0xffffffffULL && wb < 0xffffffffULL;
    p.slice_bytes = buffers_ok ? static_cast<u32>(sb) : 0;
    p.wslice_bytes = buffers_ok ? static_cast<u32>(wb) : 0;
    return buffers_ok;
}

template<class C, class In, class Out, class W>
int run(const ProjectDesc& d, hipStream_t stream, std::string& err, const Scratch* scratch) {
    InsertParams<C, In, W> p;
    const bool buffers_ok = fill_params(d, p);
    if (p.n_slices == 0) return NOA_HIP_OK;
    const KernelChoice kc = kernel_choice(d);
    const C cutoff = std::max(static_cast<C>(d.cutoff), C(0));
    const bool nearest = d.interp == NOA_HIP_INTERP_NEAREST;

    constexpr bool single = sizeof(typename vt<In>::real) == 4 && sizeof(W) == 4;
    if constexpr (single) {
        // An inverse scaling without Ewald curvature leaves fz (the culling) as it is and scales the 2D frequency.
        const bool flat = p.ews[0] == 0 && p.ews[1] == 0;
        if ((!p.general || flat) && !kc.force_general && !p.in_is_value && !p.win_is_value && buffers_ok &&
            window_size(d.interp) == 2) {
            int device = 0;
            (void)hipGetDevice(&device);
            const Plan* planp = nullptr;
            // Plans are kept per slice count rounded up to m * 2^k (8 <= m < 16, at least 64; items are clamped to
            // the actual count in the kernel), so that callers with varying batch sizes reuse a few plans and at most
            // 1/8 of a plan's slice range is empty.
            i32 step = 8;
            while (p.n_slices > 16 * step) step *= 2;
            const i32 plan_slices = std::max<i32>(64, (p.n_slices + step - 1) / step * step);
            // The slab kernel: Mat33<f32> rotations (or Quaternion<f32>, through their matrix: the coordinates are
            // noa's quaternion rotation up to rounding, within the insertion's tolerance), LINEAR, one power-of-two size n >= 32 for the slices, the
            // volume and the target, cutoff <= 0.5, a window of at most ~1.15 samples (at most 4 candidates per column)
            // that the degree-9 polynomial fits, unit x strides and one row pitch for the slices and their weights.
            bool slab = false;
            SlabWindow sw{};
            if constexpr (std::is_same_v<C, float>) {
                const i64 nn = d.target[0];
                const bool cubic = d.target[1] == nn && d.target[2] == nn && p.grid[0] == nn && p.grid[1] == nn &&
                                   p.grid[2] == nn && p.slice_shape[0] == nn && p.slice_shape[1] == nn;
                if (!kc.force_wave && !p.general && d.interp == NOA_HIP_INTERP_LINEAR && cubic && p.pow2 && nn >= 32 &&
                    (p.rot.kind == NOA_HIP_MAT33_F32 || p.rot.kind == NOA_HIP_QUAT_F32) && cutoff <= C(0.5) && p.is_x == 1 &&
                    (!p.win_present || (p.wis_x == 1 && p.wis_y == p.is_y)) && p.slice_bytes < 0x7ffffff0u &&
                    p.wslice_bytes < 0x7ffffff0u) {
                    sw = fit_slab_window(static_cast<float>(p.fftfreq_sinc), static_cast<float>(p.fftfreq_blackman),
                                         static_cast<i32>(nn));
                    slab = sw.ok && sw.fbT <= 1.15f;
                }
            }
            // Expected passing slices per work item: the table's value, else 512 for the slab kernel (12 resident
            // waves per CU: 4.6-4.7 ms against 5.0 ms at 256 for 10,000 slices at 256^3) and 256 for the one-wave
            // kernel.
            // The default adapts to the call's work (make_plan); the allocation-free entry point keeps the fixed
            // value, whose partial sums fit its scratch.
            const double cap = slab ? 512.0 : 256.0;
            const double list_target = kc.list_target > 0 ? kc.list_target : (scratch == nullptr ? -cap : cap);
            const int st = make_plan(device, p.grid, d.target, static_cast<double>(cutoff), p.vol_centered,
                                     plan_slices, static_cast<double>(p.fftfreq_blackman), p.bd, list_target,
                                     scratch == nullptr, planp, err, kc.order,
                                     scratch == nullptr && kc.order != 3 ? kc.max_span : 0);
            if (st != NOA_HIP_OK) return st;
            // Pose-adaptive plan (item order 3; not for the allocation-free entry point): the passing slices of each
            // active sub-brick are counted on the device, and the segments are cut from the counts.
            Plan pose;
            if (kc.order == 3 && scratch == nullptr && !planp->h_bricks.empty()) {
                Plan& shared = const_cast<Plan&>(*planp);
                {
                    std::lock_guard<std::mutex> lock(g_plan_mutex);
                    if (!shared.bricks &&
                        (hipMalloc(&shared.bricks, shared.h_bricks.size() * sizeof(i32)) != hipSuccess ||
                         hipMemcpy(shared.bricks, shared.h_bricks.data(), shared.h_bricks.size() * sizeof(i32),
                                   hipMemcpyHostToDevice) != hipSuccess)) {
                        err = "insert: could not upload the active sub-bricks";
                        return NOA_HIP_ERROR_RUNTIME;
                    }
                }
                const i32 nbk = static_cast<i32>(shared.h_bricks.size());
                i32* dcounts = nullptr;
                if (hipMallocAsync(&dcounts, static_cast<size_t>(nbk) * CNT_BINS * sizeof(i32), stream) != hipSuccess) {
                    err = "insert: could not allocate the pass counts";
                    return NOA_HIP_ERROR_RUNTIME;
                }
                CountParams<C> cp{};
                cp.rot = p.rot;
                cp.bricks = shared.bricks;
                cp.counts = dcounts;
                cp.n_slices = p.n_slices;
                cp.n_bricks = nbk;
                cp.bin_w = count_bin_width(p.n_slices);
                constexpr i32 RANGES = 8;
                cp.range_w = 64 * std::max<i32>(1, ((p.n_slices + 63) / 64 + RANGES - 1) / RANGES);
                const i64 n_counts = static_cast<i64>(nbk) * CNT_BINS;
                hipLaunchKernelGGL(zero_i32_kernel, dim3(static_cast<u32>((n_counts + 255) / 256)), dim3(256), 0, stream,
                                   dcounts, n_counts);
                for (int a = 0; a < 3; ++a) {
                    cp.grid[a] = p.grid[a];
                    cp.nb[a] = p.nb[a];
                    cp.bd[a] = p.bd[a];
                    cp.target[a] = p.target[a];
                }
                cp.fb = p.fftfreq_blackman;
                cp.vol_centered = p.vol_centered;
                hipLaunchKernelGGL((insert_count_kernel<C>),
                                   dim3(static_cast<u32>((nbk + CNT_BRICKS - 1) / CNT_BRICKS),
                                        static_cast<u32>((p.n_slices + cp.range_w - 1) / cp.range_w)),
                                   dim3(64), 0, stream, cp);
                std::vector<i32> counts(static_cast<size_t>(nbk) * CNT_BINS);
                const bool ok = hipMemcpyAsync(counts.data(), dcounts, counts.size() * sizeof(i32), hipMemcpyDeviceToHost,
                                               stream) == hipSuccess &&
                                hipStreamSynchronize(stream) == hipSuccess;
                (void)hipFreeAsync(dcounts, stream);
                if (!ok) { err = "insert: could not read the pass counts"; return NOA_HIP_ERROR_RUNTIME; }
                // The model plan stays when no sub-brick needs more than twice its segments.
                bool skewed = false;
                for (i32 j = 0; j < nbk && !skewed; ++j) {
                    i64 total = 0;
                    for (i32 b = 0; b < CNT_BINS; ++b) total += counts[static_cast<size_t>(j) * CNT_BINS + b];
                    skewed = static_cast<double>(total) > 2.0 * shared.target * shared.h_nseg[static_cast<size_t>(j)];
                }
                if (skewed) make_pose_plan(shared.h_bricks, counts, p.n_slices, shared.target, pose);
            }
            if (pose.n_items > 0) {
                if (hipMallocAsync(&pose.items, pose.h_items.size() * sizeof(int4), stream) != hipSuccess ||
                    hipMemcpyAsync(pose.items, pose.h_items.data(), pose.h_items.size() * sizeof(int4),
                                   hipMemcpyHostToDevice, stream) != hipSuccess ||
                    (!pose.h_reduce.empty() &&
                     (hipMallocAsync(&pose.reduce, pose.h_reduce.size() * sizeof(int4), stream) != hipSuccess ||
                      hipMemcpyAsync(pose.reduce, pose.h_reduce.data(), pose.h_reduce.size() * sizeof(int4),
                                     hipMemcpyHostToDevice, stream) != hipSuccess))) {
                    err = "insert: could not upload the pose-adaptive plan";
                    return NOA_HIP_ERROR_RUNTIME;
                }
                pose.owned = true;
                planp = &pose;
            }
            const Plan& plan = *planp;
            // The pose-adaptive plan's device arrays are released after this call's kernels (stream-ordered).
            struct PoseRelease {
                Plan& pl;
                hipStream_t st;
                ~PoseRelease() {
                    if (!pl.owned) return;
                    if (pl.items) (void)hipFreeAsync(pl.items, st);
                    if (pl.reduce) (void)hipFreeAsync(pl.reduce, st);
                }
            } pose_release{pose, stream};
            if (plan.n_items == 0) return NOA_HIP_OK;
            const size_t partial_bytes = static_cast<size_t>(plan.n_partials) * p.bd[0] * p.bd[1] * p.apitch * 3 *
                                         sizeof(float);
            const size_t record_bytes = static_cast<size_t>(p.n_slices) * 4 * sizeof(float4);
            void* partials = nullptr;
            void* records = nullptr;     // slab kernel only
            if (scratch) {
                // Allocation-free: plan, records and partial sums in the caller-independent scratch, in that order.
                auto round = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
                const size_t items_bytes = round(plan.h_items.size() * sizeof(int4));
                const size_t reduce_bytes = round(plan.h_reduce.size() * sizeof(int4));
                const size_t need = items_bytes + reduce_bytes + round(record_bytes) + round(partial_bytes);
                if (need > scratch->bytes) {
                    err = "insert (no allocation): this call needs " + std::to_string(need) +
                          " bytes of scratch, more than the " + std::to_string(scratch->bytes) +
                          "-byte maximum; insert fewer slices per call";
                    return NOA_HIP_ERROR_UNSUPPORTED;
                }
                char* at = scratch->base;
                int4* items = reinterpret_cast<int4*>(at);
                at += items_bytes;
                int4* reduce = plan.h_reduce.empty() ? nullptr : reinterpret_cast<int4*>(at);
                at += reduce_bytes;
                records = at;
                at += round(record_bytes);
                partials = plan.n_partials > 0 ? at : nullptr;
                p.items = items;
                p.reduce = reduce;
                if (hipMemcpyAsync(items, plan.h_items.data(), plan.h_items.size() * sizeof(int4),
                                   hipMemcpyHostToDevice, stream) != hipSuccess ||
                    (reduce && hipMemcpyAsync(reduce, plan.h_reduce.data(), plan.h_reduce.size() * sizeof(int4),
                                              hipMemcpyHostToDevice, stream) != hipSuccess)) {
                    err = "insert (no allocation): could not copy the work plan to the scratch";
                    return NOA_HIP_ERROR_RUNTIME;
                }
            } else {
                p.items = plan.items;
                p.reduce = plan.reduce;
                if (plan.n_partials > 0 && hipMallocAsync(&partials, partial_bytes, stream) != hipSuccess) {
                    err = "insert: could not allocate the partial sums";
                    return NOA_HIP_ERROR_RUNTIME;
                }
            }
            p.partials = static_cast<float*>(partials);
            const dim3 grid_items(static_cast<unsigned>(plan.n_items));
            // The compile-time configurations: both layouts centred or both not, and the options HOT fixes.
            const bool quat = p.rot.kind == NOA_HIP_QUAT_F32 || p.rot.kind == NOA_HIP_QUAT_F64;
            const bool hot_opts = p.pow2 && !quat && !p.general && p.win_present && p.has_wvol && p.wpoly &&
                                  p.vol_centered == p.slices_centered && p.is_x == 1 && p.wis_x == 1 &&
                                  p.is_y == p.wis_y && p.slice_bytes < 0x7ffffff0u && p.wslice_bytes < 0x7ffffff0u;
            const int hot = hot_opts ? (p.vol_centered ? 2 : 1) : 0;
            if (slab) {
                if (!scratch && hipMallocAsync(&records, record_bytes, stream) != hipSuccess) {
                    if (partials) (void)hipFreeAsync(partials, stream);
                    err = "insert: could not allocate the slice records";
                    return NOA_HIP_ERROR_RUNTIME;
                }
                if constexpr (std::is_same_v<C, float>) {
                    hipLaunchKernelGGL(slab_records_kernel, dim3(static_cast<unsigned>((p.n_slices + 255) / 256)),
                                       dim3(256), 0, stream, p.rot, p.n_slices, sw.fbT, static_cast<float4*>(records));
                    SlabParams<In, W> sp{};
                    sp.in = p.in; sp.is_b = p.is_b;
                    sp.win = p.win_present ? p.win : nullptr; sp.wis_b = p.wis_b;
                    sp.pitch = p.is_y;
                    sp.slice_bytes = p.slice_bytes; sp.wslice_bytes = p.wslice_bytes;
                    sp.vol = p.vol; sp.vs0 = p.vs0; sp.vs1 = p.vs1; sp.vs2 = p.vs2;
                    sp.wvol = p.has_wvol ? p.wvol : nullptr; sp.wvs0 = p.wvs0; sp.wvs1 = p.wvs1; sp.wvs2 = p.wvs2;
                    sp.n = p.grid[0];
                    for (int i = 0; i < 3; ++i) sp.nb[i] = p.nb[i];
                    sp.items = p.items;
                    sp.partials = p.partials;
                    sp.rec = static_cast<const float4*>(records);
                    sp.n_slices = p.n_slices;
                    sp.fbT = sw.fbT;
                    sp.tk = sw.tk;
                    for (int k = 0; k <= SLAB_NWIN; ++k) sp.wa[k] = sw.a[k];
                    sp.cut2 = static_cast<float>(p.cutoff_sqd) * static_cast<float>(p.grid[0]) * static_cast<float>(p.grid[0]);
                    sp.vol_centered = p.vol_centered;
                    if (kc.paired) {
                        if (p.slices_centered)
                            hipLaunchKernelGGL((insert_slab_kernel<In, Out, W, true, true>), grid_items, dim3(64), 0, stream, sp);
                        else
                            hipLaunchKernelGGL((insert_slab_kernel<In, Out, W, false, true>), grid_items, dim3(64), 0, stream, sp);
                    } else {
                        if (p.slices_centered)
                            hipLaunchKernelGGL((insert_slab_kernel<In, Out, W, true, false>), grid_items, dim3(64), 0, stream, sp);
                        else
                            hipLaunchKernelGGL((insert_slab_kernel<In, Out, W, false, false>), grid_items, dim3(64), 0, stream, sp);
                    }
                }
            } else {
                if (nearest) {
                    hipLaunchKernelGGL((insert_wave_kernel<C, In, Out, W, NOA_HIP_INTERP_NEAREST>), grid_items, dim3(64), 0, stream, p);
                } else if constexpr (std::is_same_v<C, float>) {
                    if (hot == 1) hipLaunchKernelGGL((insert_wave_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR, 1>), grid_items, dim3(64), 0, stream, p);
                    else if (hot == 2) hipLaunchKernelGGL((insert_wave_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR, 2>), grid_items, dim3(64), 0, stream, p);
                    else hipLaunchKernelGGL((insert_wave_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR>), grid_items, dim3(64), 0, stream, p);
                } else {
                    hipLaunchKernelGGL((insert_wave_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR>), grid_items, dim3(64), 0, stream, p);
                }
            }
            if (plan.n_reduce > 0)
                hipLaunchKernelGGL((insert_reduce_kernel<C, In, Out, W>), dim3(plan.n_reduce), dim3(NT_REDUCE), 0, stream, p);
            hipError_t e = hipGetLastError();
            for (void* buf : {partials, records}) {
                if (!buf || scratch) continue;
                const hipError_t fe = hipFreeAsync(buf, stream);
                if (e == hipSuccess) e = fe;
            }
            if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
            return NOA_HIP_OK;
        }
    }

    const i64 n_vox = static_cast<i64>(p.grid[0]) * p.grid[1] * p.wr;
    const dim3 grid(static_cast<unsigned>((n_vox + 255) / 256)), block(256);
    p.interp_mode = d.interp;
    auto launch = [&](auto interp_tag) {
        constexpr int I = decltype(interp_tag)::value;
        if (p.general) hipLaunchKernelGGL((insert_general_kernel<C, In, Out, W, I, true>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((insert_general_kernel<C, In, Out, W, I, false>), grid, block, 0, stream, p);
    };
    if (nearest) launch(std::integral_constant<int, NOA_HIP_INTERP_NEAREST>{});
    else if (d.interp == NOA_HIP_INTERP_LINEAR) launch(std::integral_constant<int, NOA_HIP_INTERP_LINEAR>{});
    else launch(std::integral_constant<int, INTERP_WINDOW>{});
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class C>
int dispatch(const ProjectDesc& d, hipStream_t stream, std::string& err, const Scratch* sc) {
    const int in = d.input.dtype, out = d.output.dtype;
    if (in == NOA_HIP_C32 && out == NOA_HIP_C32) return run<C, float2, float2, float>(d, stream, err, sc);
    if (in == NOA_HIP_C32 && out == NOA_HIP_F32) return run<C, float2, float, float>(d, stream, err, sc);
    if (in == NOA_HIP_F32 && out == NOA_HIP_F32) return run<C, float, float, float>(d, stream, err, sc);
    if (in == NOA_HIP_C64 && out == NOA_HIP_C64) return run<C, double2, double2, double>(d, stream, err, sc);
    if (in == NOA_HIP_C64 && out == NOA_HIP_F64) return run<C, double2, double, double>(d, stream, err, sc);
    if (in == NOA_HIP_F64 && out == NOA_HIP_F64) return run<C, double, double, double>(d, stream, err, sc);
    err = "insert: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

// Scratch of the allocation-free entry point: one array per GPU, resident for the life of the process. Host code
// never uses the symbol itself as an address: the device address of each GPU's copy is looked up once, with that
// GPU current, and kept.
__device__ __align__(256) unsigned char scratch[NOA_HIP_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mutex;
std::vector<void*> g_scratch_ptr;   // per device ordinal; null until looked up

int scratch_address(int device, void*& ptr, std::string& err) {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (static_cast<int>(g_scratch_ptr.size()) <= device) g_scratch_ptr.resize(static_cast<size_t>(device) + 1, nullptr);
    if (!g_scratch_ptr[static_cast<size_t>(device)]) {
        void* found = nullptr;
        const hipError_t e = hipGetSymbolAddress(&found, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !found) {
            (void)hipGetLastError();
            err = std::string("insert (no allocation): could not find the scratch array on device ") +
                  std::to_string(device) + ": " + hipGetErrorString(e);
            return NOA_HIP_ERROR_RUNTIME;
        }
        g_scratch_ptr[static_cast<size_t>(device)] = found;
    }
    ptr = g_scratch_ptr[static_cast<size_t>(device)];
    return NOA_HIP_OK;
}

} // namespace

int insert_batch_prepare(const ProjectDesc* descs, i32 n_jobs, void* image, i64 capacity, i64& bytes,
                         std::string& err) {
    using P = InsertParams<float, float2, float>;
    auto round = [](i64 b) { return (b + 255) & ~static_cast<i64>(255); };
    std::vector<P> params(static_cast<size_t>(n_jobs));
    struct Item { float work; int4 item; };
    std::vector<Item> items;
    int device = 0;
    (void)hipGetDevice(&device);
    for (i32 k = 0; k < n_jobs; ++k) {
        const ProjectDesc& d = descs[k];
        const std::string at = "insert batch, call " + std::to_string(k) + ": ";
        const bool weights_ok = !d.input_weight.present() || (d.input_weight.dtype == NOA_HIP_F32 && !d.input_weight.is_value);
        const bool wvol_ok = !d.output_weight.data || d.output_weight.dtype == NOA_HIP_F32;
        if (d.input.dtype != NOA_HIP_C32 || d.input.is_value || d.output.dtype != NOA_HIP_C32 || !weights_ok ||
            !wvol_ok || d.coord_f64 || d.general || d.interp != NOA_HIP_INTERP_LINEAR) {
            err = at + "only complex FP32 slices (arrays) into complex FP32 volumes, FP32 weights, LINEAR "
                       "interpolation and single-precision transforms without scaling or Ewald sphere are batched";
            return NOA_HIP_ERROR_UNSUPPORTED;
        }
        P& p = params[static_cast<size_t>(k)];
        if (!fill_params(d, p)) { err = at + "the slices exceed the 4 GiB of a buffer load"; return NOA_HIP_ERROR_UNSUPPORTED; }
        if (p.n_slices == 0) continue;
        i32 step = 8;
        while (p.n_slices > 16 * step) step *= 2;
        const i32 plan_slices = std::max<i32>(64, (p.n_slices + step - 1) / step * step);
        const Plan* plan = nullptr;
        // One segment per sub-brick (no partial sums): list target above any slice count.
        const int st = make_plan(device, p.grid, d.target, static_cast<double>(std::max(static_cast<float>(d.cutoff), 0.f)),
                                 p.vol_centered, plan_slices, static_cast<double>(p.fftfreq_blackman), p.bd, 1e30, false,
                                 plan, err);
        if (st != NOA_HIP_OK) return st;
        for (size_t i = 0; i < plan->h_items.size(); ++i) {
            int4 it = plan->h_items[i];
            it.w = k;
            items.push_back({plan->h_work[i], it});
        }
    }
    // Largest expected work first, across the calls.
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.work > b.work; });
    const i64 jobs_offset = round(static_cast<i64>(sizeof(InsertBatchHeader)));
    const i64 items_offset = round(jobs_offset + static_cast<i64>(sizeof(P)) * n_jobs);
    bytes = items_offset + static_cast<i64>(items.size()) * static_cast<i64>(sizeof(int4));
    if (!image || capacity < bytes) {
        err = "insert batch: the image needs " + std::to_string(bytes) + " bytes";
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    if (items_offset > 0x7fffffffLL || items.size() > 0x7fffffffULL) { err = "insert batch: too many calls or items"; return NOA_HIP_ERROR_UNSUPPORTED; }
    unsigned char* out = static_cast<unsigned char*>(image);
    std::memset(out, 0, static_cast<size_t>(items_offset));
    InsertBatchHeader h{};
    h.n_items = static_cast<i32>(items.size());
    h.n_jobs = n_jobs;
    h.jobs_offset = static_cast<i32>(jobs_offset);
    h.items_offset = static_cast<i32>(items_offset);
    std::memcpy(out, &h, sizeof(h));
    if (n_jobs > 0) std::memcpy(out + jobs_offset, params.data(), sizeof(P) * static_cast<size_t>(n_jobs));
    int4* dst = reinterpret_cast<int4*>(out + items_offset);
    for (size_t i = 0; i < items.size(); ++i) dst[i] = items[i].item;
    return NOA_HIP_OK;
}

int insert_batch_launch(const void* image, i32 grid, hipStream_t stream, std::string& err) {
    if (grid < 1) { err = "insert batch: the grid needs at least one workgroup"; return NOA_HIP_ERROR_INVALID_ARGUMENT; }
    hipLaunchKernelGGL(insert_batch_kernel, dim3(static_cast<u32>(grid)), dim3(64), 0, stream,
                       static_cast<const unsigned char*>(image));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

int launch_insert(const ProjectDesc& d, hipStream_t stream, std::string& err, bool no_alloc) {
    if (!no_alloc) return d.coord_f64 ? dispatch<double>(d, stream, err, nullptr) : dispatch<float>(d, stream, err, nullptr);
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess) {
        err = "insert (no allocation): no current device";
        return NOA_HIP_ERROR_RUNTIME;
    }
    void* base = nullptr;
    const int st = scratch_address(device, base, err);
    if (st != NOA_HIP_OK) return st;
    const Scratch sc{static_cast<char*>(base), NOA_HIP_NOALLOC_SCRATCH_BYTES};
    return d.coord_f64 ? dispatch<double>(d, stream, err, &sc) : dispatch<float>(d, stream, err, &sc);
}

} // namespace noa_hip
