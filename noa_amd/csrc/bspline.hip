// noa::xform::cubic_bspline_prefilter (src/noa/xform/CubicBSplinePrefilter.hpp:24): the recursive filter that
// turns samples into cubic B-spline coefficients, so that CUBIC_BSPLINE interpolation passes through the samples.
// Per line (core/CubicBSplinePrefilter.hpp:22-107): a causal recursion c[i] = L * s[i] + z * c[i-1] started from
// L * (s[0] + sum_{k < min(12, n)} z^(k+1) s[k]), then an anticausal one c[i] = z * (c[i+1] - c[i]) started from
// z / (z - 1) * c[n-1], with z = sqrt(3) - 2 and L = (1 - z)(1 - 1/z). Lines along x, then y, then z of the
// dimensions reordered to the output's rightmost order (cpu/CubicBSplinePrefilter.hpp:129-158).
//
// The recursion is sequential along a line, so one lane owns one line, with noa's expressions in noa's order
// (bit-identical to the CPU path). What differs from a lane-per-line loop over memory is the access pattern:
//  - lines along y and z: the 64 lanes of a wave own 64 neighbouring x, so every step of the recursion is one
//    coalesced 64-element access; steps are loaded in batches of 16, the next batch in flight while the current
//    one's dependent arithmetic runs;
//  - lines along x (the contiguous axis): a wave owns 64 rows and walks them in 32-column chunks staged through
//    LDS, loaded and stored as whole row segments; the lanes then run their row's recursion out of LDS while the
//    next chunk's loads are in flight.
// Each pass reads and writes the array once; the anticausal half re-reads the causal half's output, which the
// column passes do straight from the L2/last-level cache and the row pass from LDS for the last chunk.
#include "common.hpp"
#include "kernels.hpp"

namespace noa_hip {

template<class T> __device__ __forceinline__ T vsub(T a, T b);
template<> __device__ __forceinline__ float vsub(float a, float b) { return a - b; }
template<> __device__ __forceinline__ double vsub(double a, double b) { return a - b; }
template<> __device__ __forceinline__ float2 vsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
template<> __device__ __forceinline__ double2 vsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// The pole and gain in the value's real precision, as constants of that precision (core/CubicBSplinePrefilter.hpp:
// 37-38); the anticausal factor z / (z - 1) (:104).
template<class R> struct BSplineConst {
    static constexpr R POLE = static_cast<R>(-0.2679491924311228);
    static constexpr R LAMBDA = (1 - POLE) * (1 - 1 / POLE);
    static constexpr R ANTI = POLE / (POLE - 1);
};

constexpr int UNROLL = 16;         // steps per batch in the column passes (two batches in flight)

template<class T>
struct LinesParams {
    const T* in; T* out;
    i64 in_s[3], out_s[3];         // strides of the three outer (line-index) dimensions
    i64 in_step, out_step;         // stride along the line
    i32 n1, n2;                    // sizes of the two inner outer dimensions (n2 fastest)
    i64 lines;                     // number of lines
    i32 n;                         // samples per line
};

// Rows: one wave per 64 lines, chunks of CHUNK columns through LDS. Block = one wave. The next chunk's global
// loads are issued into registers before the current chunk's recursion runs, so their latency overlaps it.
template<class T>
__global__ __launch_bounds__(64) void bspline_rows_kernel(const LinesParams<T> p) {
    using R = typename vt<T>::real;
    using K = BSplineConst<R>;
    // Columns per LDS chunk (64 columns of 4-byte values measured no faster at 256^3 and 512^3).
    constexpr i32 CHUNK = 32;
    constexpr i32 RPS = 64 / CHUNK;   // rows moved per load step
    __shared__ T tile[64][CHUNK + 1];
    __shared__ i64 ibase[64], obase[64];
    const i32 lane = static_cast<i32>(threadIdx.x);
    const i64 first = static_cast<i64>(blockIdx.x) * 64;
    const i32 nrows = static_cast<i32>(min<i64>(64, p.lines - first));
    {
        const i64 line = first + lane;
        i64 ib = 0, ob = 0;
        if (lane < nrows) {
            const i64 c = line % p.n2, t = line / p.n2;
            const i64 b = t % p.n1, a = t / p.n1;
            ib = a * p.in_s[0] + b * p.in_s[1] + c * p.in_s[2];
            ob = a * p.out_s[0] + b * p.out_s[1] + c * p.out_s[2];
        }
        ibase[lane] = ib;
        obase[lane] = ob;
    }
    __syncthreads();

    // Cooperative chunk copies: step k moves rows RPS*k .. RPS*k + RPS-1, CHUNK lanes along each row.
    const i32 col = lane % CHUNK, rsub = lane / CHUNK;
    // 16-byte values would need 128 registers for the prefetch: they are loaded straight into LDS instead.
    constexpr bool PREFETCH = sizeof(T) <= 8;
    T regs[PREFETCH ? CHUNK : 1];
    auto load_regs = [&](const T* src, const i64* base, i64 step, i32 c0, i32 w) {
#pragma unroll
        for (i32 k = 0; k < CHUNK; ++k) {
            const i32 r = RPS * k + rsub;
            if (col < w && r < nrows) regs[k] = src[base[r] + static_cast<i64>(c0 + col) * step];
        }
    };
    auto regs_to_tile = [&](i32 w) {
#pragma unroll
        for (i32 k = 0; k < CHUNK; ++k) {
            const i32 r = RPS * k + rsub;
            if (col < w && r < nrows) tile[r][col] = regs[k];
        }
    };
    auto load_tile = [&](const T* src, const i64* base, i64 step, i32 c0, i32 w) {
#pragma unroll
        for (i32 k = 0; k < CHUNK; ++k) {
            const i32 r = RPS * k + rsub;
            if (col < w && r < nrows) tile[r][col] = src[base[r] + static_cast<i64>(c0 + col) * step];
        }
    };
    auto store_chunk = [&](i32 c0, i32 w) {
#pragma unroll
        for (i32 k = 0; k < CHUNK; ++k) {
            const i32 r = RPS * k + rsub;
            if (col < w && r < nrows) p.out[obase[r] + static_cast<i64>(c0 + col) * p.out_step] = tile[r][col];
        }
    };

    const i32 n = p.n;
    const i32 last = ((n - 1) / CHUNK) * CHUNK;
    T prev = vzero<T>();
    // Causal pass (filter(), :60-73): the input's first chunk holds the min(12, n) samples of the initialisation.
    if constexpr (PREFETCH) load_regs(p.in, ibase, p.in_step, 0, min(CHUNK, n));
    for (i32 c0 = 0; c0 < n; c0 += CHUNK) {
        const i32 w = min(CHUNK, n - c0);
        if constexpr (PREFETCH) regs_to_tile(w);
        else load_tile(p.in, ibase, p.in_step, c0, w);
        __syncthreads();
        if constexpr (PREFETCH)
            if (c0 != last) load_regs(p.in, ibase, p.in_step, c0 + CHUNK, min(CHUNK, n - c0 - CHUNK));
        if (lane < nrows) {
            T* row = tile[lane];
            i32 j = 0;
            if (c0 == 0) {
                const i32 horizon = min(12, n);
                R zn = K::POLE;
                T sum = row[0];
                for (i32 k = 0; k < horizon; ++k) {
                    sum = vadd(sum, vmul(row[k], zn));
                    zn *= K::POLE;
                }
                prev = vmul(sum, K::LAMBDA);
                row[0] = prev;
                j = 1;
            }
            for (; j < w; ++j) {
                prev = vadd(vmul(row[j], K::LAMBDA), vmul(prev, K::POLE));
                row[j] = prev;
            }
        }
        __syncthreads();
        if (c0 != last) {
            store_chunk(c0, w);
            __syncthreads();
        }
    }
    // Anticausal pass (:75-79), from the last chunk, still in LDS, back to the first.
    for (i32 c0 = last; c0 >= 0; c0 -= CHUNK) {
        const i32 w = min(CHUNK, n - c0);
        if (c0 != last) {
            if constexpr (PREFETCH) regs_to_tile(w);
            else load_tile(p.out, obase, p.out_step, c0, w);
            __syncthreads();
        }
        if constexpr (PREFETCH)
            if (c0 != 0) load_regs(p.out, obase, p.out_step, c0 - CHUNK, CHUNK);
        if (lane < nrows) {
            T* row = tile[lane];
            i32 j = w - 1;
            if (c0 == last) {
                prev = vmul(row[j], K::ANTI);
                row[j] = prev;
                --j;
            }
            for (; j >= 0; --j) {
                prev = vmul(vsub(prev, row[j]), K::POLE);
                row[j] = prev;
            }
        }
        __syncthreads();
        store_chunk(c0, w);
        __syncthreads();
    }
}

// Columns and pages, in place on the output: one lane per line, neighbouring lanes on neighbouring x. Steps are
// loaded UNROLL at a time, the next batch issued before the current one's dependent arithmetic.
template<class T>
__global__ __launch_bounds__(256) void bspline_lines_kernel(const LinesParams<T> p) {
    using R = typename vt<T>::real;
    using K = BSplineConst<R>;
    const i64 line = static_cast<i64>(blockIdx.x) * 256 + threadIdx.x;
    if (line >= p.lines) return;
    const i64 c = line % p.n2, t = line / p.n2;
    const i64 b = t % p.n1, a = t / p.n1;
    T* const base = p.out + (a * p.out_s[0] + b * p.out_s[1] + c * p.out_s[2]);
    const i64 st = p.out_step;
    const i32 n = p.n;

    // Causal initialisation (:87-100) and recursion (:67-73).
    const i32 horizon = min(12, n);
    R zn = K::POLE;
    T sum = base[0];
    for (i32 k = 0; k < horizon; ++k) {
        sum = vadd(sum, vmul(base[k * st], zn));
        zn *= K::POLE;
    }
    T prev = vmul(sum, K::LAMBDA);
    base[0] = prev;
    T cur[UNROLL], nxt[UNROLL];
#pragma unroll
    for (i32 j = 0; j < UNROLL; ++j)
        if (1 + j < n) cur[j] = base[(1 + j) * st];
    for (i32 i = 1; i < n; i += UNROLL) {
        const i32 in = i + UNROLL;
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j)
            if (in + j < n) nxt[j] = base[(in + j) * st];
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j) {
            if (i + j < n) {
                prev = vadd(vmul(cur[j], K::LAMBDA), vmul(prev, K::POLE));
                base[(i + j) * st] = prev;
            }
        }
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j) cur[j] = nxt[j];
    }
    // Anticausal initialisation (:102-106) from c[n-1], which `prev` holds, and recursion (:75-79).
    prev = vmul(prev, K::ANTI);
    base[(n - 1) * st] = prev;
#pragma unroll
    for (i32 j = 0; j < UNROLL; ++j)
        if (n - 2 - j >= 0) cur[j] = base[(n - 2 - j) * st];
    for (i32 i = n - 2; i >= 0; i -= UNROLL) {
        const i32 in = i - UNROLL;
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j)
            if (in - j >= 0) nxt[j] = base[(in - j) * st];
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j) {
            if (i - j >= 0) {
                prev = vmul(vsub(prev, cur[j]), K::POLE);
                base[(i - j) * st] = prev;
            }
        }
#pragma unroll
        for (i32 j = 0; j < UNROLL; ++j) cur[j] = nxt[j];
    }
}

namespace {

// Launch of one pass over `lines` lines of `n` samples; the outer dimensions are (lines / (n1 * n2), n1, n2).
template<class T>
int pass(bool rows, const T* in, const i64 in_s[3], i64 in_step, T* out, const i64 out_s[3], i64 out_step,
         i64 n0, i64 n1, i64 n2, i64 n, hipStream_t stream, std::string& err) {
    LinesParams<T> p{};
    p.in = in;
    p.out = out;
    for (int i = 0; i < 3; ++i) { p.in_s[i] = in_s[i]; p.out_s[i] = out_s[i]; }
    p.in_step = in_step;
    p.out_step = out_step;
    p.n1 = static_cast<i32>(n1);
    p.n2 = static_cast<i32>(n2);
    p.lines = n0 * n1 * n2;
    p.n = static_cast<i32>(n);
    if (p.lines == 0 || n == 0) return NOA_HIP_OK;
    if (rows) {
        const i64 blocks = (p.lines + 63) / 64;
        hipLaunchKernelGGL((bspline_rows_kernel<T>), dim3(static_cast<u32>(blocks)), dim3(64), 0, stream, p);
    } else {
        const i64 blocks = (p.lines + 255) / 256;
        hipLaunchKernelGGL((bspline_lines_kernel<T>), dim3(static_cast<u32>(blocks)), dim3(256), 0, stream, p);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { err = std::string("kernel launch failed: ") + hipGetErrorString(e); return NOA_HIP_ERROR_RUNTIME; }
    return NOA_HIP_OK;
}

template<class T>
int run(const BSplineDesc& d, hipStream_t stream, std::string& err) {
    const T* in = static_cast<const T*>(d.input.data);
    T* out = static_cast<T*>(d.output.data);
    const i64* is = d.input.s;
    const i64* os = d.output.s;
    const i64* sh = d.shape;
    int status;
    // Rows (x), from the input; then columns (y) and pages (z), in place.
    const i64 ri[3] = {is[0], is[1], is[2]}, ro[3] = {os[0], os[1], os[2]};
    if ((status = pass<T>(true, in, ri, is[3], out, ro, os[3], sh[0], sh[1], sh[2], sh[3], stream, err))) return status;
    if (d.ndim >= 2) {
        const i64 co[3] = {os[0], os[1], os[3]};
        if ((status = pass<T>(false, out, co, os[2], out, co, os[2], sh[0], sh[1], sh[3], sh[2], stream, err))) return status;
    }
    if (d.ndim == 3) {
        const i64 po[3] = {os[0], os[2], os[3]};
        if ((status = pass<T>(false, out, po, os[1], out, po, os[1], sh[0], sh[2], sh[3], sh[1], stream, err))) return status;
    }
    return NOA_HIP_OK;
}

} // namespace

int launch_bspline_prefilter(const BSplineDesc& d, hipStream_t stream, std::string& err) {
    switch (d.output.dtype) {
        case NOA_HIP_F32: return run<float>(d, stream, err);
        case NOA_HIP_F64: return run<double>(d, stream, err);
        case NOA_HIP_C32: return run<float2>(d, stream, err);
        case NOA_HIP_C64: return run<double2>(d, stream, err);
        default: err = "cubic_bspline_prefilter: f32, f64, c32 or c64 arrays are expected"; return NOA_HIP_ERROR_UNSUPPORTED;
    }
}

} // namespace noa_hip
