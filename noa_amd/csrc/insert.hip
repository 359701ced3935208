// Fourier insertion of central slices by 2D interpolation with a windowed sinc along the slice normal.
//
// Replaces FourierInsertInterpolate (src/noa/xform/FourierProject.hpp:392-539) launched by
// launch_insert_central_slices_3d (:1319-1394).
//
// Each voxel of the stored half-volume is owned by one lane, which sums the contributions of the slices in
// increasing slice order and then adds the sum to the volume (read-modify-write, no atomics): the summation
// order is the reference's, so the result is deterministic and bit-comparable with the CPU path up to the
// sin/cos of the windowed sinc.
//
// A slice only contributes to voxels within its windowed-sinc slab |fz| <= fftfreq_blackman, i.e. about 0.6 %
// of the (voxel, slice) pairs at 256^3. The volume is cut into bricks of 4x8x8 voxels (one 256-lane workgroup
// per brick); a first kernel tests every (brick, slice) pair against a conservative slab bound (the brick's
// bounding sphere) and writes, per brick, the ordered list of slices that may touch it. The gather kernel
// then walks only that list. Skipped slices would have added exactly zero, so the result is unchanged.
#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <mutex>
#include <unordered_map>

namespace noa_hip {

constexpr i32 BZ = 4, BY = 8, BX = 8;   // brick = 256 voxels = one workgroup

template<class C, class In, class W>
struct InsertParams {
    const In* in; i64 is_b; i32 is_y, is_x; bool in_is_value; In in_value;
    const W* win; i64 wis_b; i32 wis_y, wis_x; bool win_present, win_is_value; W win_value;
    void* vol; i32 vs0, vs1, vs2;
    W* wvol; i32 wvs0, wvs1, wvs2; bool has_wvol;
    i32 grid[3];          // logical D, H, W
    i32 wr;               // W/2+1
    i32 bricks[3];        // bricks along z, y, x
    C target[3];
    C slice_f[2];
    i32 slice_shape[2];   // logical H, W of the slices
    i32 n_slices;
    C cutoff_sqd, fftfreq_sinc, fftfreq_blackman;
    C ews[2];
    XformArg<C> rot, scale;
    bool general, slices_centered, vol_centered;
    // Brick culling (list mode): per brick, `counts[brick]` slice indices at lists[brick * n_slices ...].
    const i32* lists; const i32* counts;
};

// Voxel coordinates of lane `t` of brick `brick`.
__device__ __forceinline__ void brick_voxel(i32 brick, i32 t, const i32 nb[3], i32& z, i32& y, i32& x) {
    const i32 bx = brick % nb[2];
    const i32 r = brick / nb[2];
    const i32 by = r % nb[1];
    const i32 bz = r / nb[1];
    z = bz * BZ + t / (BY * BX);
    y = by * BY + (t / BX) % BY;
    x = bx * BX + t % BX;
}

// Conservative test of every (brick, slice) pair: does the slab |fz| <= fftfreq_blackman of the slice
// intersect the bounding sphere of the brick (in normalised frequencies)? One lane per brick; the slices are
// walked in order so that each brick's list is sorted.
template<class C>
__global__ __launch_bounds__(256) void insert_cull_kernel(const InsertParams<C, float, float> p, i32* lists,
                                                          i32* counts) {
    const i32 n_bricks = p.bricks[0] * p.bricks[1] * p.bricks[2];
    const i32 brick = static_cast<i32>(blockIdx.x) * 256 + static_cast<i32>(threadIdx.x);
    if (brick >= n_bricks) return;
    // Brick centre and radius in normalised frequencies (double: the bound must not be tighter than exact).
    i32 z0, y0, x0;
    brick_voxel(brick, 0, p.bricks, z0, y0, x0);
    const double hz = 0.5 * (BZ - 1), hy = 0.5 * (BY - 1), hx = 0.5 * (BX - 1);
    // Frequencies of the brick's voxels along z and y may wrap (non-centred layouts): use the min/max of the
    // frequencies actually covered, which are contiguous inside a brick except across the wrap point.
    double lo[3], hi[3];
    const i32 starts[3] = {z0, y0, x0};
    const i32 sizes[3] = {BZ, BY, BX};
    for (int a = 0; a < 3; ++a) {
        lo[a] = 1e30; hi[a] = -1e30;
        for (int k = 0; k < sizes[a]; ++k) {
            const i32 idx = starts[a] + k;
            const i32 n = p.grid[a];
            if (a < 2 && idx >= n) continue;
            if (a == 2 && idx >= p.wr) continue;
            const i32 fr = a < 2 ? index2frequency(p.vol_centered, idx, n) : idx;
            const double f = static_cast<double>(fr) / static_cast<double>(p.target[a]);
            lo[a] = fmin(lo[a], f); hi[a] = fmax(hi[a], f);
        }
    }
    (void)hz; (void)hy; (void)hx;
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) { counts[brick] = 0; return; }
    double c[3], rad2 = 0, closest2 = 0;
    for (int a = 0; a < 3; ++a) {
        c[a] = 0.5 * (lo[a] + hi[a]);
        const double h = 0.5 * (hi[a] - lo[a]);
        rad2 += h * h;
        const double d = lo[a] > 0 ? lo[a] : (hi[a] < 0 ? -hi[a] : 0.0);
        closest2 += d * d;
    }
    i32 n = 0;
    if (closest2 <= static_cast<double>(p.cutoff_sqd) * (1.0 + 1e-6) + 1e-12) {
        const double rad = sqrt(rad2);
        // Bound on |fz| error between float arithmetic and this double test: generous relative slack.
        const double slab = static_cast<double>(p.fftfreq_blackman) * (1.0 + 1e-4) + rad + 1e-6;
        for (i32 i = 0; i < p.n_slices; ++i) {
            const Rot<C> r = load_rotation(p.rot, i);
            double fz;
            if (!r.quat) {
                fz = (static_cast<double>(r.m[0]) * c[0] + static_cast<double>(r.m[1]) * c[1]) + static_cast<double>(r.m[2]) * c[2];
            } else {
                vec3<double> v{c[0], c[1], c[2]};
                fz = quat_rotate<double>(r.m[0], r.m[1], r.m[2], r.m[3], v).z;
            }
            // The row norm of a rotation is 1; allow for non-normalised input with the actual norm.
            double norm;
            if (!r.quat) {
                norm = sqrt(static_cast<double>(r.m[0]) * r.m[0] + static_cast<double>(r.m[1]) * r.m[1] +
                            static_cast<double>(r.m[2]) * r.m[2]);
            } else {
                norm = static_cast<double>(r.m[0]) * r.m[0] + static_cast<double>(r.m[1]) * r.m[1] +
                       static_cast<double>(r.m[2]) * r.m[2] + static_cast<double>(r.m[3]) * r.m[3];
            }
            if (fabs(fz) <= slab * fmax(norm, 1.0))
                lists[static_cast<i64>(brick) * p.n_slices + n++] = i;
        }
    }
    counts[brick] = n;
}

template<class C, class In, class Out, class W, int INTERP, bool GENERAL>
__global__ __launch_bounds__(256) void insert_kernel(const InsertParams<C, In, W> p) {
    using RIn = typename vt<In>::real;
    const i32 brick = static_cast<i32>(xcd_swizzle(blockIdx.x, gridDim.x));
    i32 z, y, x;
    brick_voxel(brick, static_cast<i32>(threadIdx.x), p.bricks, z, y, x);
    if (z >= p.grid[0] || y >= p.grid[1] || x >= p.wr) return;

    // operator() (:478-519)
    vec3<C> f;
    f.z = static_cast<C>(index2frequency(p.vol_centered, z, p.grid[0])) / p.target[0];
    f.y = static_cast<C>(index2frequency(p.vol_centered, y, p.grid[1])) / p.target[1];
    f.x = static_cast<C>(x) / p.target[2];
    if (dot3(f) > p.cutoff_sqd) return;

    In value = vzero<In>();
    W weights{};
    const i32 count = p.counts[brick];
    const i32* list = p.lists + static_cast<i64>(brick) * p.n_slices;
    for (i32 k = 0; k < count; ++k) {
        const i32 i = list[k];
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
                                             p.slices_centered, py, px, p.in_is_value, p.in_value);
            value = vadd(value, vmul(v, static_cast<RIn>(window)));
            if (p.has_wvol) {
                if (p.win_present) {
                    const W* wbase = p.win + static_cast<i64>(i) * p.wis_b;
                    const W wv = interp2<INTERP, W>(wbase, p.wis_y, p.wis_x, p.slice_shape[0], p.slice_shape[1],
                                                    p.slices_centered, py, px, p.win_is_value, p.win_value);
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

// Per-device scratch for the brick lists, grown on demand and kept for the process lifetime. Calls that
// overlap on one device from different host threads are serialised by the mutex only up to the launch; the
// kernels of two overlapping calls on different streams would share the buffer, so such calls take a private
// allocation instead.
struct Scratch {
    void* ptr{};
    size_t bytes{};
};
std::mutex g_scratch_mutex;
std::unordered_map<int, Scratch> g_scratch;

template<class C, class In, class Out, class W>
int run(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    InsertParams<C, In, W> p{};
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
    for (int i = 0; i < 3; ++i) {
        p.grid[i] = static_cast<i32>(d.out_shape[i + 1]);
        p.target[i] = static_cast<C>(d.target[i]);
    }
    p.wr = p.grid[2] / 2 + 1;
    p.bricks[0] = (p.grid[0] + BZ - 1) / BZ;
    p.bricks[1] = (p.grid[1] + BY - 1) / BY;
    p.bricks[2] = (p.wr + BX - 1) / BX;
    p.slice_f[0] = static_cast<C>(d.in_shape[2]);
    p.slice_f[1] = static_cast<C>(d.in_shape[3]);
    p.slice_shape[0] = static_cast<i32>(d.in_shape[2]);
    p.slice_shape[1] = static_cast<i32>(d.in_shape[3]);
    p.n_slices = static_cast<i32>(d.in_shape[0]);
    C cutoff = std::max(static_cast<C>(d.cutoff), C(0));
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

    const i64 n_bricks = static_cast<i64>(p.bricks[0]) * p.bricks[1] * p.bricks[2];
    if (n_bricks == 0 || p.n_slices == 0) return NOA_HIP_OK;

    // Brick lists: n_bricks counts + n_bricks * n_slices indices.
    const size_t need = static_cast<size_t>(n_bricks) * (static_cast<size_t>(p.n_slices) + 1) * sizeof(i32);
    int device = 0;
    (void)hipGetDevice(&device);
    void* scratch = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        Scratch& s = g_scratch[device];
        if (s.bytes < need) {
            if (s.ptr) {
                // The previous buffer may still be used by queued kernels of an earlier call.
                (void)hipDeviceSynchronize();
                (void)hipFree(s.ptr);
                s.ptr = nullptr; s.bytes = 0;
            }
            if (hipMalloc(&s.ptr, need) != hipSuccess) {
                err = "insert: could not allocate the brick lists";
                return NOA_HIP_ERROR_RUNTIME;
            }
            s.bytes = need;
        }
        scratch = s.ptr;
    }
    i32* counts = static_cast<i32*>(scratch);
    i32* lists = counts + n_bricks;
    p.counts = counts;
    p.lists = lists;

    InsertParams<C, float, float> q{};
    q.grid[0] = p.grid[0]; q.grid[1] = p.grid[1]; q.grid[2] = p.grid[2]; q.wr = p.wr;
    for (int i = 0; i < 3; ++i) { q.bricks[i] = p.bricks[i]; q.target[i] = p.target[i]; }
    q.n_slices = p.n_slices; q.cutoff_sqd = p.cutoff_sqd; q.fftfreq_blackman = p.fftfreq_blackman;
    q.rot = p.rot; q.vol_centered = p.vol_centered;
    const unsigned cull_blocks = static_cast<unsigned>((n_bricks + 255) / 256);
    if (p.general) {
        // Scaling/EWS bend or stretch the slab: cull nothing (every brick gets the full list).
        // The list then holds 0..n_slices-1; it is written by the same kernel with an infinite slab.
        q.fftfreq_blackman = static_cast<C>(1e30);
    }
    hipLaunchKernelGGL((insert_cull_kernel<C>), dim3(cull_blocks), dim3(256), 0, stream, q, lists, counts);

    const dim3 grid(static_cast<unsigned>(n_bricks)), block(256);
    const bool nearest = d.interp == NOA_HIP_INTERP_NEAREST;
    if (p.general) {
        if (nearest) hipLaunchKernelGGL((insert_kernel<C, In, Out, W, NOA_HIP_INTERP_NEAREST, true>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((insert_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR, true>), grid, block, 0, stream, p);
    } else {
        if (nearest) hipLaunchKernelGGL((insert_kernel<C, In, Out, W, NOA_HIP_INTERP_NEAREST, false>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((insert_kernel<C, In, Out, W, NOA_HIP_INTERP_LINEAR, false>), grid, block, 0, stream, p);
    }
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
    err = "insert: unsupported input/output value types";
    return NOA_HIP_ERROR_UNSUPPORTED;
}

} // namespace

int launch_insert(const ProjectDesc& d, hipStream_t stream, std::string& err) {
    return d.coord_f64 ? dispatch<double>(d, stream, err) : dispatch<float>(d, stream, err);
}

} // namespace noa_hip
