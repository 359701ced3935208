// Volume tiles of 16^3 sample positions, each visited by the slices whose plane crosses it: the enumeration shared by
// the tiled extraction and the tiled rasterisation, and their work plan.
//
// Positions are the floor of a slice pixel's (flipped) coordinates in samples. With cutoff <= 0.5 every pixel inside
// the cutoff has its floor within [lo - 1, hi + 1] along z and y and [0, n/2] along x, which the tiles cover. For a
// tile and a slice whose plane crosses it, tile_candidates lists the pixels whose position may lie in the tile --
// directly, or mirrored by the x < 0 flip -- row by row: lanes 0-31 take rows of the direct set and lanes 32-63 rows of
// the mirrored one, each computes its row's pixel range from the tile's box (with a margin far above the rounding), a
// prefix sum packs the candidates into the wave's LDS list, and the caller's fetch/store run on them, two per lane at
// a time. The caller keeps a candidate only where its own floor lies in the tile, so each pixel inside the cutoff is
// handled by exactly one tile.
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace noa_hip {

constexpr i32 XT = 16;                  // tile extent in floor positions along each axis
constexpr i32 XTT = XT + 1;             // LDS extent along each axis (the trilinear taps reach floor + 1)
constexpr i32 XT_LIST = 768;            // candidate pixels per wave and round
constexpr i32 XT_WAVES = 4;             // waves per work item

// Candidate word: (row frequency + 512) | pixel x << 10 | mirrored << 31.
__device__ __forceinline__ u32 tile_word(i32 fy, i32 u, bool mirrored) {
    return static_cast<u32>(fy + 512) | (static_cast<u32>(u) << 10) | (mirrored ? 0x80000000u : 0u);
}

// Does the plane of `r` (spanned by its 2nd and 3rd columns; any matrix) cross the box of centre c and half-extent
// XT/2 widened by M? The plane is symmetric through the origin, so this also decides the mirrored box.
__device__ __forceinline__ bool tile_crossed(const Rot<float>& r, float cz, float cy, float cx, float M) {
    const float nz = r.m[4] * r.m[8] - r.m[7] * r.m[5];
    const float ny = r.m[7] * r.m[2] - r.m[1] * r.m[8];
    const float nx = r.m[1] * r.m[5] - r.m[4] * r.m[2];
    const float d = fabsf((nz * cz + ny * cy) + nx * cx);
    const float sum = fabsf(nz) + fabsf(ny) + fabsf(nx);
    return d <= (0.5f * XT) * sum + M * sum + 1e-6f;
}

// The same test for a box of EZ x 16 x 16 positions (a slab of a tile), centre c.
template<int EZ>
__device__ __forceinline__ bool slab_crossed(const Rot<float>& r, float cz, float cy, float cx, float M) {
    const float nz = r.m[4] * r.m[8] - r.m[7] * r.m[5];
    const float ny = r.m[7] * r.m[2] - r.m[1] * r.m[8];
    const float nx = r.m[1] * r.m[5] - r.m[4] * r.m[2];
    const float d = fabsf((nz * cz + ny * cy) + nx * cx);
    const float sum = fabsf(nz) + fabsf(ny) + fabsf(nx);
    return d <= (0.5f * EZ) * fabsf(nz) + (0.5f * XT) * (fabsf(ny) + fabsf(nx)) + M * sum + 1e-6f;
}

// The candidate pixels of one (box, crossing slice) pair; the box is [Z0, Z0 + EZ) x [Y0, Y0 + 16) x [X0, X0 + 16)
// (a whole tile when EZ = 16). Cand: the caller's candidate type with a `bool ok`; fetch(fy, u, mirrored, Cand&)
// evaluates one candidate, store(const Cand&) finishes it. `wl`: the wave's list of LIST words in LDS, which must
// hold a whole row (n/2 + 1 <= LIST: rows enter the list whole). A degenerate transform (collapsed plane) makes every
// pixel of the slice a candidate.
template<class Cand, int EZ = XT, int LIST = XT_LIST, class Fetch, class Store>
__device__ __forceinline__ void tile_candidates(const Rot<float>& r, i32 Z0, i32 Y0, i32 X0, i32 n, float M, i32 pixels,
                                                float inv_wr, bool slice_centered, i32 lane, u32* wl, Fetch& fetch,
                                                Store& store) {
    constexpr i32 EXT[3] = {EZ, XT, XT};
    const i32 lo = -n / 2, hi = (n - 1) / 2, hix = n / 2;
    // Plane coordinates: F = Y a + U c (a, c: the 2nd and 3rd columns); on the plane Y = g . F with g the
    // first row of (G^T G)^-1 G^T, G = [a c].
    const float a[3] = {r.m[1], r.m[4], r.m[7]}, c[3] = {r.m[2], r.m[5], r.m[8]};
    const float aa = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    const float ac = (a[0] * c[0] + a[1] * c[1]) + a[2] * c[2];
    const float cc = (c[0] * c[0] + c[1] * c[1]) + c[2] * c[2];
    const float det = aa * cc - ac * ac;
    if (!(det > 1e-6f * aa * cc)) {
        // Degenerate transform (the plane collapses): every pixel of the slice is a candidate.
        for (i32 pix = lane; pix < pixels; pix += 64) {
            const i32 yq = static_cast<i32>((static_cast<float>(pix) + 0.5f) * inv_wr);
            const i32 u = pix - yq * (hix + 1);
            const i32 fy = index2frequency(slice_centered, yq, n);
            Cand c;
            fetch(fy, u, false, c);
            store(c);
            fetch(fy, u, true, c);
            store(c);
        }
        return;
    }
    const float idet = 1.f / det;
    const float g[3] = {(cc * a[0] - ac * c[0]) * idet, (cc * a[1] - ac * c[1]) * idet,
                        (cc * a[2] - ac * c[2]) * idet};
    const bool mirrored = lane >= 32;
    // The box of this lane's set, widened by the margin: direct F in [P0, P0 + EXT), mirrored F in
    // (-P0 - EXT, -P0].
    const float P0[3] = {static_cast<float>(Z0), static_cast<float>(Y0), static_cast<float>(X0)};
    float blo[3], bhi[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        blo[i] = (mirrored ? -(P0[i] + EXT[i]) : P0[i]) - M;
        bhi[i] = (mirrored ? -P0[i] : P0[i] + EXT[i]) + M;
    }
    float ymin = 0.f, ymax = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        ymin += g[i] >= 0.f ? g[i] * blo[i] : g[i] * bhi[i];
        ymax += g[i] >= 0.f ? g[i] * bhi[i] : g[i] * blo[i];
    }
    const i32 yfirst = max(static_cast<i32>(ceilf(fmaxf(ymin, static_cast<float>(lo) - 1.f))), lo);
    const i32 ylast = min(static_cast<i32>(floorf(fminf(ymax, static_cast<float>(hi) + 1.f))), hi);
    for (i32 rb = 0;; rb += 32) {
        const i32 fy = yfirst + rb + (lane & 31);
        const bool row = fy <= ylast;
        if (__ballot(row) == 0) break;
        // Pixel range of the row: U with Y a + U c inside the box.
        float ulo = 0.f, uhi = static_cast<float>(hix);
        bool any = row;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float base = a[i] * static_cast<float>(fy);
            if (fabsf(c[i]) > 1e-12f) {
                const float t1 = (blo[i] - base) / c[i], t2 = (bhi[i] - base) / c[i];
                ulo = fmaxf(ulo, fminf(t1, t2));
                uhi = fminf(uhi, fmaxf(t1, t2));
            } else if (base < blo[i] || base > bhi[i]) {
                any = false;
            }
        }
        i32 u0 = 0, cnt = 0;
        if (any && ulo <= uhi) {
            u0 = static_cast<i32>(ceilf(ulo));
            cnt = max(0, static_cast<i32>(floorf(uhi)) - u0 + 1);
        }
        const i32 incl = wave_inclusive_sum(cnt), excl = incl - cnt;
        const i32 total = __builtin_amdgcn_readlane(incl, 63);
        for (i32 done = 0; done < total;) {
            // Whole rows into the list, as many as fit (rows are short: at most n/2 + 1 pixels).
            const bool fits = cnt > 0 && excl >= done && incl <= done + LIST;
            const unsigned long long fm = __ballot(fits);
            const i32 last = 63 - __clzll(static_cast<long long>(fm));
            const i32 upto = __builtin_amdgcn_readlane(incl, last);
            if (fits)
                for (i32 q = 0; q < cnt; ++q) wl[excl - done + q] = tile_word(fy, u0 + q, mirrored);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const i32 len = upto - done;
            for (i32 e = lane; e < len; e += 128) {
                Cand c0, c1;
                const u32 w0 = wl[e];
                fetch(static_cast<i32>(w0 & 1023u) - 512, static_cast<i32>((w0 >> 10) & 1023u), (w0 >> 31) != 0u,
                      c0);
                c1.ok = false;
                if (e + 64 < len) {
                    const u32 w1 = wl[e + 64];
                    fetch(static_cast<i32>(w1 & 1023u) - 512, static_cast<i32>((w1 >> 10) & 1023u),
                          (w1 >> 31) != 0u, c1);
                }
                store(c0);
                store(c1);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            done = upto;
        }
    }
}

// Work items of the tiled extraction for one geometry: every tile that the cutoff sphere reaches, cut into ranges
// of slices so that each item expects about `pairs` crossing slices (a plane through the origin with a uniformly
// drawn normal crosses a box of half-extent h at distance r with probability min(1, 1.5 h / r)); items sorted by
// expected work, largest first. Kept per (device, size, cutoff, slice count, pairs) for the process lifetime.
struct TilePlan {
    int4* items{};
    i32 n_items{}, ntz{}, nty{}, ntx{}, base{};
};
inline std::mutex g_tile_mutex;
inline std::map<std::tuple<int, i32, double, i32, i32>, TilePlan> g_tile_plans;

// A negative `pairs` is a default: at most -pairs expected crossing slices per item, fewer when the call's expected
// crossings would fill fewer than TILE_MIN_ITEMS items (small calls: more, shorter items keep every CU busy).
constexpr double TILE_MIN_ITEMS = 512;

inline int make_tile_plan(int device, i32 n, double cutoff, i32 n_slices, i32 pairs, TilePlan& plan, std::string& err) {
    const auto key = std::make_tuple(device, n, cutoff, n_slices, pairs);
    std::lock_guard<std::mutex> lock(g_tile_mutex);
    auto it = g_tile_plans.find(key);
    if (it != g_tile_plans.end()) {
        plan = it->second;
        return NOA_HIP_OK;
    }
    TilePlan pl;
    const i32 lo = -n / 2, hi = (n - 1) / 2;
    pl.base = lo - 1;
    pl.ntz = pl.nty = (hi + 1 - (lo - 1) + 1 + XT - 1) / XT;
    pl.ntx = (n / 2 + 1 + XT - 1) / XT;
    const double rc = cutoff * n + 1.0;
    struct Entry { double work; int4 item; };
    struct Active { i32 tile; double expected; };
    std::vector<Active> active;
    std::vector<Entry> entries;
    for (i32 tz = 0; tz < pl.ntz; ++tz)
        for (i32 ty = 0; ty < pl.nty; ++ty)
            for (i32 tx = 0; tx < pl.ntx; ++tx) {
                const double lo3[3] = {double(pl.base + tz * XT), double(pl.base + ty * XT), double(tx * XT)};
                double closest = 0, r2 = 0;
                for (int a = 0; a < 3; ++a) {
                    const double l = lo3[a], h = lo3[a] + XT;
                    const double dd = l > 0 ? l : (h < 0 ? -h : 0.0);
                    closest += dd * dd;
                    const double c = l + 0.5 * XT;
                    r2 += c * c;
                }
                if (closest > rc * rc) continue;
                const double frac = std::min(1.0, 1.5 * 0.5 * XT / std::max(std::sqrt(r2), 1e-9));
                active.push_back({(tz * pl.nty + ty) * pl.ntx + tx, frac * n_slices});
            }
    double per_item = pairs;
    if (pairs < 0) {
        double total = 0;
        for (const Active& a : active) total += a.expected;
        per_item = std::min(static_cast<double>(-pairs), std::max(8.0, total / TILE_MIN_ITEMS));
    }
    for (const Active& a : active) {
        const i32 n_seg = std::max<i32>(1, static_cast<i32>(std::ceil(a.expected / per_item)));
        const i32 len = (n_slices + n_seg - 1) / n_seg;
        for (i32 sgm = 0; sgm < n_seg; ++sgm) {
            const i32 b = sgm * len, e = std::min(n_slices, b + len);
            if (b < e) entries.push_back({a.expected / n_seg + (e - b) / 64.0, make_int4(a.tile, b, e, 0)});
        }
    }
    std::stable_sort(entries.begin(), entries.end(), [](const Entry& a, const Entry& b) { return a.work > b.work; });
    std::vector<int4> items(entries.size());
    for (size_t k = 0; k < entries.size(); ++k) items[k] = entries[k].item;
    pl.n_items = static_cast<i32>(items.size());
    if (!items.empty() &&
        (hipMalloc(&pl.items, items.size() * sizeof(int4)) != hipSuccess ||
         hipMemcpy(pl.items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice) != hipSuccess)) {
        err = "extract: could not upload the tile plan";
        return NOA_HIP_ERROR_RUNTIME;
    }
    g_tile_plans[key] = pl;
    plan = pl;
    return NOA_HIP_OK;
}

} // namespace noa_hip
