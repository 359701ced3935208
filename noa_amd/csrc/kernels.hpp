// Resolved (validated, flattened) parameters of the three operators and their launchers. The C ABI layer
// (api.cpp) checks a noa_hip_project_args, fills one of these structures and calls the launcher.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/noa_hip.h"

namespace noa_hip {

using i32 = int32_t;
using i64 = int64_t;

// A flattened strided array: slices use (batch, y, x) strides, volumes (z, y, x) strides.
struct Arr {
    void* data{};
    i64 s0{}, s1{}, s2{};
    int dtype{};          // NOA_HIP_* or NOA_HIP_DTYPE_NONE (Empty)
    bool is_value{};
    double value[2]{};
    bool present() const { return data != nullptr || is_value; }
};

struct Xform {
    const void* data{};
    int kind{};           // NOA_HIP_XFORM_NONE = Empty
    bool batched{};
    double single[9]{};
};

// Resolved per-batch 2D shifts of a phase shift (single shifts are copied from the host).
struct ShiftDesc {
    bool present{};
    const void* data{};
    bool f64{};
    bool batched{};
    double single[2]{};
    double cutoff{1};
    int mode{};           // 0: general formula, 1: no shift (copy/remap), 2: half-shape shift (+-1 pattern)
};

// Everything the kernels of one call need; coordinates are converted to the rotation precision on launch.
struct ProjectDesc {
    Arr input, input_weight, output, output_weight;
    i64 in_shape[4]{}, out_shape[4]{};   // logical BDHW
    i64 target[3]{};                     // logical DHW of the target (resolved default)
    bool src_centered{}, dst_centered{};
    int interp{};                        // NOEA_HIP_INTERP_* with _FAST erased
    Xform rot, scale;
    bool general{};                      // scaling or EWS present
    double ews_radius[2]{};
    double cutoff{}, fftfreq_sinc{}, fftfreq_blackman{};
    bool coord_f64{};
    ShiftDesc shift;                     // extraction only: phase shift of the output slices (fused)
};

// Extraction from a virtual volume of inserted central slices (insert_and_extract_central_slices_3d).
struct InsertExtractDesc {
    Arr input, input_weight, output, output_weight;   // slices: (batch, y, x) strides
    i64 in_shape[4]{}, out_shape[4]{};                 // logical BDHW
    bool src_centered{}, dst_centered{};
    int interp{};
    Xform rot_in, scale_in, rot_out, scale_out;        // inv rotation + fwd scaling (insert), fwd + inv (extract)
    bool general{};                                    // any scaling or EWS present
    double ews_radius[2]{};
    double cutoff{}, in_sinc{}, in_blackman{}, w_sinc{}, w_blackman{};
    i64 w_size{};                                      // samples of the w window (1: the 3D operator)
    bool add{}, correct{};
    bool coord_f64{};
};

// noa::signal::ctf_anisotropic over (B, 1, H, W/2+1) rfft arrays; with `output_weight`, the fused form
// (output = input * ctf, output_weight = ctf^2).
struct CtfDesc {
    Arr input, output, output_weight;    // (batch, y, x) strides; input.data == nullptr: write the CTF
    i64 shape[4]{};                      // logical BDHW
    bool src_centered{}, dst_centered{};
    bool ctf_abs{}, ctf_squared{};
    double range_start{}, range_stop{-1};
    bool endpoint{true};
    bool ctf_f64{}, ctf_batched{};
    const void* ctf_data{};              // device array of CTFAnisotropic<T> when batched
    alignas(16) unsigned char ctf_single[112]{};   // one CTFAnisotropic<T>, copied from the host
};

// A phase shift over (B, 1, H, W/2+1) rfft arrays.
struct PhaseShiftDesc {
    Arr input, output;    // (batch, y, x) strides; input.data == nullptr: write the phase shifts
    i64 b{}, h{}, w{};    // logical batch, height, width
    bool src_centered{}, dst_centered{};
    bool c64{};
    ShiftDesc shift;
};

// A strided 4D array (b, z, y, x strides) and the elementwise operators on spectra and images.
struct Arr4 {
    void* data{};
    i64 s[4]{};
    int dtype{};
};
// A CTF over 1, 2 or 3 dimensions (noa's general CTF operator), after noa's reordering of the dimensions: dims are
// the iterated dimensions, outermost first, the last one being the half dimension of a half layout.
struct CtfNdDesc {
    Arr4 input, output;        // s[0]: batch stride, s[1 + i]: stride of dims[i]; input.data == nullptr: the CTF
    int n_dims{};              // 1, 2 or 3
    i64 batch{};
    i64 dims[3]{};             // logical sizes
    bool half{}, src_centered{}, dst_centered{};
    bool isotropic{};          // CTFIsotropic objects, else CTFAnisotropic (2 dimensions)
    bool ctf_abs{}, ctf_squared{};
    double range_start{}, range_stop{-1};
    bool endpoint{true};
    bool ctf_f64{}, ctf_batched{};
    const void* ctf_data{};    // device array of CTF objects when batched
    alignas(16) unsigned char ctf_single[112]{};
};

struct RemapDesc {
    Arr4 input, output;
    i64 shape[4]{};          // logical BDHW
    bool src_centered{}, dst_centered{};
};
// FFTs (rocFFT): the logical BDHW shape (the real array's for r2c and c2r), element strides of each array in its own
// element type, the rank picked by noa (1: a row or column vector, 2, 3).
enum { FFT_R2C = 0, FFT_C2R = 1, FFT_C2C = 2 };
struct FftDesc {
    int type{};
    int sign{-1};            // -1 forward, +1 backward (r2c: forward, c2r: backward)
    bool f64{};
    void* in{}; void* out{}; // equal: in place
    i64 shape[4]{};
    i64 in_s[4]{}, out_s[4]{};
    int rank{};
    int norm{};
    bool cache_plan{true}, plan_only{}, record_workspace{};
    int device{};
};

// One chunk of MRC data: `count` elements from element `first` (BDHW order of the array's shape), between raw file
// bytes in device memory and a strided f32/f64/c32/c64 device array.
struct MrcChunk {
    void* raw{};
    int file_dtype{};        // NOA_HIP_IO_*
    bool swap{}, clamp{};
    i64 first{}, count{};
    void* array{};
    int array_dtype{};
    i64 s[4]{}, n[4]{};
};

// Cubic B-spline prefilter: dimensions already reordered to the output's rightmost order, input strides broadcast
// (0 along broadcast dimensions); ndim of the reordered shape picks the passes (x; x, y; x, y, z).
struct BSplineDesc {
    Arr4 input, output;
    i64 shape[4]{};
    int ndim{};
};

struct GriddingDesc {
    Arr4 input, output;      // input batch stride 0 when broadcast
    i64 shape[4]{};          // BDHW of the output
    bool post{};
};

// noa's blackman_window_size (FourierProject.hpp:114-137): samples of the w window, always odd.
inline i64 blackman_window_size(double fftfreq_blackman, double spectrum_size) {
    double samples = spectrum_size * fftfreq_blackman;
    if (std::abs(samples - 1.) <= 1e-6) samples = std::round(samples);
    else samples = std::ceil(samples);
    const i64 n = std::max<i64>(1, static_cast<i64>(samples));
    return 2 * n + 1 - 2;
}

// Launchers: return 0 or a NOA_HIP_ERROR_* code, with a message in `err`.
int launch_remap(const RemapDesc& d, hipStream_t stream, std::string& err);
int launch_gridding_correction(const GriddingDesc& d, hipStream_t stream, std::string& err);
int launch_phase_shift(const PhaseShiftDesc& d, hipStream_t stream, std::string& err);
int launch_extract(const ProjectDesc& d, hipStream_t stream, std::string& err);
int launch_rasterize(const ProjectDesc& d, hipStream_t stream, std::string& err);
// no_alloc: the plan, slice records and partial sums go to the library's static device scratch (no allocation;
// calls that overlap on one GPU race on it).
int launch_insert(const ProjectDesc& d, hipStream_t stream, std::string& err, bool no_alloc = false);
int launch_insert_extract(const InsertExtractDesc& d, hipStream_t stream, std::string& err);
// Batched extraction / insertion (one persistent launch for several calls); prepare writes the device image on the
// host, launch reads it from device memory.
i64 volume_blocks_bytes(const int64_t rfft_shape[4]);
int volume_blocks(const void* volume, i64 s_z, i64 s_y, const int64_t rfft_shape[4], void* blocks, hipStream_t stream,
                  std::string& err);
i64 extract_batch_bytes(i32 n_jobs);
int extract_batch_prepare(const ProjectDesc* descs, i32 n_jobs, void* image, i64 capacity, i32 ppt, i64& bytes,
                          std::string& err);
int extract_batch_launch(const void* image, i32 grid, hipStream_t stream, std::string& err);
int insert_batch_prepare(const ProjectDesc* descs, i32 n_jobs, void* image, i64 capacity, i64& bytes,
                         std::string& err);
int insert_batch_launch(const void* image, i32 grid, hipStream_t stream, std::string& err);
// The launch tuning table: values[] for a key (zeros where nothing matches); true if an entry matched.
bool tuning_lookup(int op, i64 box, int remap, int interp, int dtype, i32 values[4]);
int launch_ctf(const CtfDesc& d, hipStream_t stream, std::string& err);
int launch_ctf_nd(const CtfNdDesc& d, hipStream_t stream, std::string& err);
int launch_bspline_prefilter(const BSplineDesc& d, hipStream_t stream, std::string& err);
int launch_fft(const FftDesc& d, hipStream_t stream, std::string& err);
int launch_mrc_decode(const MrcChunk& c, hipStream_t stream, std::string& err);
int launch_mrc_encode(const MrcChunk& c, hipStream_t stream, std::string& err);
int mrc_read_header(const char* path, noa_hip_mrc_header& h, std::string& err);
int mrc_read(const char* path, const noa_hip_view& out, bool clamp, hipStream_t stream, std::string& err);
int mrc_write(const char* path, const noa_hip_view& in, const double* spacing, int dtype, const double* stats,
              bool clamp, hipStream_t stream, std::string& err);
int launch_strided_copy(const void* src, const i64 ss[4], void* dst, const i64 ds[4], const i64 shape[4],
                        bool f64, hipStream_t stream, std::string& err);
int fft_clear_cache(int device);
int fft_set_cache_limit(int count, int device);
int fft_cache_limit(int device);
int fft_cache_size(int device);
i64 fft_workspace_left_to_allocate(int device);
int fft_set_workspace(int device, void* buffer, i64 bytes, std::string& err);

} // namespace noa_hip
