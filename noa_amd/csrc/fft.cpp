// noa::fft::r2c / c2r / c2c (src/noa/fft/Transform.hpp:160-349) on rocFFT, with noa's plan cache
// (src/noa/fft/cuda/Plan.cpp:120-370): per host thread and per device, at most `limit` plans (4 by default),
// newest first, the oldest evicted; plans created with record_and_share_workspace get their work buffer later, one
// buffer shared by all of them (allocated by the next ordinary transform, or given by set_workspace).
//
// What differs from noa's CUDA path is where the normalisation happens: noa scales the output in a second
// elementwise pass (Transform.hpp:146-158); here the factor is the plan's rocFFT scale factor, applied by the
// transform's last kernel as it stores the result, so the output is written once. The factor is noa's, computed in
// the transform's precision: 1 / N (or 1 / sqrt(N) for ORTHO) with N the logical D * H * W.
//
// rocFFT plans are one direction and one placement each (cuFFT plans hold both directions), so the cache key
// holds the transform type, the sign, the placement and the scale factor next to the shape and strides.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <algorithm>
#include <memory>
#include <mutex>
#include <string>
#include <utility>

#include "../../include/noa_hip.h"
#include "kernels.hpp"

namespace noa_hip {
namespace {

const char* rocfft_status_str(rocfft_status s) {
    switch (s) {
        case rocfft_status_success: return "success";
        case rocfft_status_failure: return "failure";
        case rocfft_status_invalid_arg_value: return "invalid argument value";
        case rocfft_status_invalid_dimensions: return "invalid dimensions";
        case rocfft_status_invalid_array_type: return "invalid array type";
        case rocfft_status_invalid_strides: return "invalid strides";
        case rocfft_status_invalid_distance: return "invalid distance";
        case rocfft_status_invalid_offset: return "invalid offset";
        case rocfft_status_invalid_work_buffer: return "invalid work buffer";
        default: return "unknown rocFFT status";
    }
}

struct DeviceFree {
    void operator()(void* p) const noexcept { if (p) (void)hipFree(p); }
};
using Workspace = std::shared_ptr<void>;

// One rocFFT plan, its execution info and its work buffer (if the plan needs one).
struct Plan {
    rocfft_plan plan{};
    rocfft_execution_info info{};
    size_t work_bytes{};
    bool has_workspace{};
    Workspace workspace;
    Plan() = default;
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    ~Plan() {
        if (info) (void)rocfft_execution_info_destroy(info);
        if (plan) (void)rocfft_plan_destroy(plan);
    }
};

struct Cache {
    std::deque<std::pair<std::string, std::shared_ptr<Plan>>> queue;   // newest first
    size_t limit{4};
    size_t workspace_size{};    // largest work buffer still to allocate (record_and_share_workspace plans)

    int clear() {
        const int n = static_cast<int>(queue.size());
        queue.clear();
        workspace_size = 0;
        return n;
    }
    std::shared_ptr<Plan> find(const std::string& key) const {
        for (const auto& [k, p]: queue)
            if (k == key) return p;
        return {};
    }
    bool has(const std::string& key) const { return static_cast<bool>(find(key)); }
    void reserve_one() {
        if (!queue.empty() && queue.size() >= limit) queue.pop_back();
    }
    void pop() {
        if (!queue.empty()) queue.pop_back();
    }
};

constexpr int MAX_DEVICES = 64;

// One cache per host thread and device (Plan.cpp:258-272). The caches are never destroyed: destroying rocFFT
// plans and freeing their buffers from thread-exit destructors can run after the HIP runtime has been torn down at
// process exit, so a thread's plans live until the process ends (or until clear_cache).
Cache& cache_of(int device) {
    thread_local Cache* caches[MAX_DEVICES] = {};
    Cache*& c = caches[device];
    if (!c) c = new Cache();
    return *c;
}

void setup_once() {
    static std::once_flag flag;
    std::call_once(flag, [] { (void)rocfft_setup(); });
}

// Allocates a work buffer, releasing the oldest cached plans (and their buffers) when the allocation fails
// (Plan.cpp:200-214).
Workspace allocate_workspace(Cache& c, size_t bytes, std::string& err) {
    const int trials = static_cast<int>(c.queue.size()) + 1;
    for (int i = 0; i < trials; ++i) {
        if (i > 0) c.pop();
        void* p = nullptr;
        if (hipMalloc(&p, bytes) == hipSuccess) return Workspace(p, DeviceFree{});
        (void)hipGetLastError();
    }
    char buf[128];
    std::snprintf(buf, sizeof(buf), "Could not allocate %zu bytes for the rocFFT workspace", bytes);
    err = buf;
    return {};
}

int attach_workspace(Plan& p, const Workspace& ws, std::string& err) {
    const rocfft_status s = rocfft_execution_info_set_work_buffer(p.info, ws.get(), p.work_bytes);
    if (s != rocfft_status_success) {
        err = std::string("rocfft_execution_info_set_work_buffer: ") + rocfft_status_str(s);
        return NOA_HIP_ERROR_RUNTIME;
    }
    p.workspace = ws;
    p.has_workspace = true;
    return NOA_HIP_OK;
}

// The work buffer of the plans created with record_and_share_workspace, shared by all of them
// (Plan.cpp:195-226).
int allocate_and_share(Cache& c, std::string& err) {
    if (c.workspace_size == 0) return NOA_HIP_OK;
    Workspace ws = allocate_workspace(c, c.workspace_size, err);
    if (!ws) return NOA_HIP_ERROR_RUNTIME;
    for (auto& [k, p]: c.queue) {
        if (!p->has_workspace) {
            const int st = attach_workspace(*p, ws, err);
            if (st != NOA_HIP_OK) return st;
        }
    }
    c.workspace_size = 0;
    return NOA_HIP_OK;
}

// noa's normalisation factor (Transform.hpp:146-158) in the transform's precision, or 1 when this direction is
// not scaled.
template<class R>
double norm_factor(const FftDesc& d) {
    const bool forward = d.sign < 0;
    const bool scaled = (forward && (d.norm == NOA_HIP_NORM_FORWARD || d.norm == NOA_HIP_NORM_ORTHO)) ||
                        (!forward && (d.norm == NOA_HIP_NORM_BACKWARD || d.norm == NOA_HIP_NORM_ORTHO));
    if (!scaled) return 1.0;
    const R count = static_cast<R>(d.shape[1] * d.shape[2] * d.shape[3]);
    const R scale = d.norm == NOA_HIP_NORM_ORTHO ? std::sqrt(count) : count;
    return static_cast<double>(1 / scale);
}

std::string make_key(const FftDesc& d, int rank, const size_t* lengths, const size_t* is, size_t idist,
                     const size_t* os, size_t odist, size_t batch, bool inplace, double factor) {
    std::string key;
    char buf[96];
    std::snprintf(buf, sizeof(buf), "%d:%d:%d:%d:%d:", d.type, d.sign, d.f64 ? 1 : 0, inplace ? 1 : 0, rank);
    key += buf;
    for (int i = 0; i < rank; ++i) {
        std::snprintf(buf, sizeof(buf), "%zu,%zu,%zu;", lengths[i], is[i], os[i]);
        key += buf;
    }
    unsigned long long fbits;
    std::memcpy(&fbits, &factor, sizeof(fbits));
    std::snprintf(buf, sizeof(buf), "%zu:%zu:%zu:%llx", idist, odist, batch, fbits);
    key += buf;
    return key;
}

int create_plan(const FftDesc& d, int rank, const size_t* lengths, const size_t* is, size_t idist,
                const size_t* os, size_t odist, size_t batch, bool inplace, double factor, Plan& p,
                std::string& err) {
    rocfft_plan_description desc = nullptr;
    rocfft_status s = rocfft_plan_description_create(&desc);
    if (s != rocfft_status_success) { err = std::string("rocfft_plan_description_create: ") + rocfft_status_str(s); return NOA_HIP_ERROR_RUNTIME; }
    rocfft_array_type in_t, out_t;
    rocfft_transform_type tt;
    if (d.type == FFT_R2C) {
        in_t = rocfft_array_type_real; out_t = rocfft_array_type_hermitian_interleaved; tt = rocfft_transform_type_real_forward;
    } else if (d.type == FFT_C2R) {
        in_t = rocfft_array_type_hermitian_interleaved; out_t = rocfft_array_type_real; tt = rocfft_transform_type_real_inverse;
    } else {
        in_t = out_t = rocfft_array_type_complex_interleaved;
        tt = d.sign < 0 ? rocfft_transform_type_complex_forward : rocfft_transform_type_complex_inverse;
    }
    s = rocfft_plan_description_set_data_layout(desc, in_t, out_t, nullptr, nullptr, static_cast<size_t>(rank), is, idist,
                                                static_cast<size_t>(rank), os, odist);
    if (s == rocfft_status_success && factor != 1.0) s = rocfft_plan_description_set_scale_factor(desc, factor);
    if (s != rocfft_status_success) {
        (void)rocfft_plan_description_destroy(desc);
        err = std::string("Failed to describe the rocFFT plan: ") + rocfft_status_str(s);
        return NOA_HIP_ERROR_INVALID_ARGUMENT;
    }
    s = rocfft_plan_create(&p.plan, inplace ? rocfft_placement_inplace : rocfft_placement_notinplace, tt,
                           d.f64 ? rocfft_precision_double : rocfft_precision_single, static_cast<size_t>(rank),
                           lengths, batch, desc);
    (void)rocfft_plan_description_destroy(desc);
    if (s != rocfft_status_success) {
        p.plan = nullptr;
        err = std::string("Failed to make the plan. ") + rocfft_status_str(s);
        return NOA_HIP_ERROR_RUNTIME;
    }
    s = rocfft_plan_get_work_buffer_size(p.plan, &p.work_bytes);
    if (s == rocfft_status_success) s = rocfft_execution_info_create(&p.info);
    if (s != rocfft_status_success) {
        err = std::string("rocFFT plan set-up: ") + rocfft_status_str(s);
        return NOA_HIP_ERROR_RUNTIME;
    }
    p.has_workspace = p.work_bytes == 0;
    return NOA_HIP_OK;
}

} // namespace

int launch_fft(const FftDesc& d, hipStream_t stream, std::string& err) {
    setup_once();
    // Rank and lengths, fastest dimension first (get_plan, Plan.cpp:411-470): D, H, W of the logical shape, a
    // column vector turned into a row vector.
    i64 sh[3] = {d.shape[1], d.shape[2], d.shape[3]};
    i64 is4[4] = {d.in_s[0], d.in_s[1], d.in_s[2], d.in_s[3]};
    i64 os4[4] = {d.out_s[0], d.out_s[1], d.out_s[2], d.out_s[3]};
    const int rank = d.rank;
    if (rank == 1 && sh[2] == 1) {
        std::swap(sh[1], sh[2]);
        std::swap(is4[2], is4[3]);
        std::swap(os4[2], os4[3]);
    }
    size_t lengths[3], is[3], os[3];
    for (int i = 0; i < rank; ++i) {
        lengths[i] = static_cast<size_t>(sh[2 - i]);
        is[i] = static_cast<size_t>(is4[3 - i]);
        os[i] = static_cast<size_t>(os4[3 - i]);
    }
    const size_t batch = static_cast<size_t>(d.shape[0]);
    size_t idist = static_cast<size_t>(d.in_s[0]), odist = static_cast<size_t>(d.out_s[0]);
    const bool inplace = d.in == d.out;

    // rocFFT's even-length real transforms read and write the real data as pairs of reals: a real operand with a
    // non-unit x stride, or an odd stride along another transformed dimension or an odd batch distance, comes out
    // shifted (measured on MI355X with rocFFT 1.0.36). Such an operand is staged through a contiguous buffer.
    const bool real_in = d.type == FFT_R2C, real_out = d.type == FFT_C2R;
    bool stage = false;
    if ((real_in || real_out) && !inplace && lengths[0] % 2 == 0) {
        const size_t* rs = real_in ? is : os;
        const size_t rdist = real_in ? idist : odist;
        stage = rs[0] != 1 || (batch > 1 && rdist % 2 != 0);
        for (int i = 1; i < rank; ++i) stage = stage || (lengths[i] > 1 && rs[i] % 2 != 0);
    }
    if (stage) {
        size_t* rs = real_in ? is : os;
        rs[0] = 1;
        for (int i = 1; i < rank; ++i) rs[i] = rs[i - 1] * lengths[i - 1];
        (real_in ? idist : odist) = rs[rank - 1] * lengths[rank - 1];
    }
    const double factor = d.f64 ? norm_factor<double>(d) : norm_factor<float>(d);
    const std::string key = make_key(d, rank, lengths, is, idist, os, odist, batch, inplace, factor);

    Cache& c = cache_of(d.device);
    int st;
    // Plans created earlier with record_and_share_workspace get their shared buffer now (Plan.cpp:313-314).
    if (!d.record_workspace && (st = allocate_and_share(c, err)) != NOA_HIP_OK) return st;

    std::shared_ptr<Plan> plan;
    if (d.plan_only || d.record_workspace) {
        if (!d.cache_plan || c.limit == 0 || c.has(key)) return NOA_HIP_OK;
    } else {
        plan = c.find(key);
    }
    if (!plan) {
        if (d.cache_plan) c.reserve_one();
        // A failed creation releases a cached plan and tries again (Plan.cpp:354-366).
        const int trials = static_cast<int>(c.queue.size()) + 1;
        st = NOA_HIP_ERROR_RUNTIME;
        for (int i = 0; i < trials; ++i) {
            if (i > 0) c.pop();
            plan = std::make_shared<Plan>();
            st = create_plan(d, rank, lengths, is, idist, os, odist, batch, inplace, factor, *plan, err);
            if (st == NOA_HIP_OK) break;
        }
        if (st != NOA_HIP_OK) return st;
        if (!plan->has_workspace && !d.record_workspace) {
            Workspace ws = allocate_workspace(c, plan->work_bytes, err);
            if (!ws) return NOA_HIP_ERROR_RUNTIME;
            if ((st = attach_workspace(*plan, ws, err)) != NOA_HIP_OK) return st;
        }
        // noa caches every new plan while the cache has room for one (Plan.cpp:368-371), cache_plan or not.
        if (c.limit > 0) {
            c.queue.emplace_front(key, plan);
            if (d.record_workspace && !plan->has_workspace) c.workspace_size = std::max(c.workspace_size, plan->work_bytes);
        }
        if (d.plan_only || d.record_workspace) return NOA_HIP_OK;
    }
    if (!plan->has_workspace) { err = "Trying to run a plan without a workspace"; return NOA_HIP_ERROR_RUNTIME; }

    // Staging buffer (stream-ordered): the contiguous copy of the real operand.
    void* tmp = nullptr;
    const i64 contiguous[4] = {d.shape[1] * d.shape[2] * d.shape[3], d.shape[2] * d.shape[3], d.shape[3], 1};
    if (stage) {
        const size_t bytes = static_cast<size_t>(d.shape[0] * contiguous[0]) * (d.f64 ? 8 : 4);
        if (hipMallocAsync(&tmp, bytes, stream) != hipSuccess) {
            (void)hipGetLastError();
            err = "Could not allocate the staging buffer of a strided real FFT operand";
            return NOA_HIP_ERROR_RUNTIME;
        }
        if (real_in && (st = launch_strided_copy(d.in, d.in_s, tmp, contiguous, d.shape, d.f64, stream, err)) != NOA_HIP_OK) {
            (void)hipFreeAsync(tmp, stream);
            return st;
        }
    }
    rocfft_status s = rocfft_execution_info_set_stream(plan->info, stream);
    if (s == rocfft_status_success) {
        void* in_buf[1] = {stage && real_in ? tmp : d.in};
        void* out_buf[1] = {stage && real_out ? tmp : d.out};
        s = rocfft_execute(plan->plan, in_buf, inplace ? nullptr : out_buf, plan->info);
    }
    if (s == rocfft_status_success && stage && real_out)
        st = launch_strided_copy(tmp, contiguous, d.out, d.out_s, d.shape, d.f64, stream, err);
    else
        st = NOA_HIP_OK;
    if (tmp) (void)hipFreeAsync(tmp, stream);
    if (s != rocfft_status_success) {
        err = std::string("rocfft_execute: ") + rocfft_status_str(s);
        return NOA_HIP_ERROR_RUNTIME;
    }
    return st;
}

int fft_clear_cache(int device) {
    if (device < 0 || device >= MAX_DEVICES) return 0;
    return cache_of(device).clear();
}

int fft_set_cache_limit(int count, int device) {
    if (device < 0 || device >= MAX_DEVICES || count < 0) return 0;
    Cache& c = cache_of(device);
    int removed = 0;
    while (c.queue.size() > static_cast<size_t>(count)) {
        c.queue.pop_back();
        ++removed;
    }
    c.limit = static_cast<size_t>(count);
    return removed;
}

int fft_cache_limit(int device) {
    if (device < 0 || device >= MAX_DEVICES) return -1;
    return static_cast<int>(cache_of(device).limit);
}

int fft_cache_size(int device) {
    if (device < 0 || device >= MAX_DEVICES) return -1;
    return static_cast<int>(cache_of(device).queue.size());
}

i64 fft_workspace_left_to_allocate(int device) {
    if (device < 0 || device >= MAX_DEVICES) return 0;
    return static_cast<i64>(cache_of(device).workspace_size);
}

// set_workspace (Plan.cpp:228-250): a caller buffer at least as large as the recorded requirement becomes the
// work buffer of every cached plan still without one; returns how many plans took it.
int fft_set_workspace(int device, void* buffer, i64 bytes, std::string& err) {
    if (device < 0 || device >= MAX_DEVICES) return 0;
    Cache& c = cache_of(device);
    if (c.workspace_size == 0 || !buffer || bytes < 0 || static_cast<size_t>(bytes) < c.workspace_size) return 0;
    const Workspace ws(buffer, [](void*) {});   // borrowed: the caller keeps it alive while the plans use it
    int count = 0;
    for (auto& [k, p]: c.queue) {
        if (!p->has_workspace) {
            if (attach_workspace(*p, ws, err) != NOA_HIP_OK) return -1;
            ++count;
        }
    }
    c.workspace_size = 0;
    return count;
}

} // namespace noa_hip
