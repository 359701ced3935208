/*
 * noa_hip.h -- C ABI of the MI355X (gfx950) central-slice library.
 *
 * Plain C: pointers, sizes and enums only. No HIP, torch or C++ types appear in the signatures, so the
 * library can be bound from C++, Python (ctypes), or any FFI.
 *
 * Each entry point replaces the GPU branch of one noa public function, i.e. the code between the
 * parameter check and the CUDA iwise launch:
 *   noa_hip_rasterize_central_slices_3d  <- noa::xform::rasterize_central_slices_3d
 *                                           (src/noa/xform/FourierProject.hpp:1653-1692, launcher :1260-1317)
 *   noa_hip_insert_central_slices_3d     <- noa::xform::insert_central_slices_3d
 *                                           (src/noa/xform/FourierProject.hpp:1770-1815, launcher :1319-1394)
 *   noa_hip_extract_central_slices_3d    <- noa::xform::extract_central_slices_3d
 *                                           (src/noa/xform/FourierProject.hpp:1869-1914, launcher :1396-1487)
 *
 * Conventions (same as noa):
 *   - Shapes are BDHW. Array shapes are rfft shapes, i.e. logical_shape.rfft() = (B, D, H, W/2+1)
 *     (src/noa/runtime/core/Shape.hpp:380). Strides are in elements, not bytes.
 *   - Slices are (B, 1, H, W/2+1) stacks; a volume is (1, D, H, W/2+1) (FourierProject.hpp:1149-1159).
 *   - Complex values are interleaved {real, imag} pairs (src/noa/base/Complex.hpp:25-29).
 *   - Mat33 elements are row-major 3x3 with a 48-byte stride for f32 (alignas(16) Mat, base/Mat.hpp:177)
 *     and 80 bytes for f64; Quaternions are {z, y, x, w} (xform/core/Quaternion.hpp:15-23);
 *     Mat22 are row-major 2x2 (16 B for f32, 32 B for f64).
 *   - Rasterisation and extraction take forward rotations; interpolated insertion takes inverse rotations
 *     (FourierProject.hpp:1750, :1768).
 *
 * Semantics:
 *   - Rasterisation and insertion ACCUMULATE into the volume (and its weights); extraction OVERWRITES the
 *     slices and writes 0 beyond the cutoff (FourierProject.hpp:334-340, :516-518, :654-656).
 *   - All calls are asynchronous on `stream` (a hipStream_t; NULL = the null stream of the current device).
 *     Buffers are borrowed: the caller keeps them alive until the stream has passed the call.
 *   - Every call returns NOA_HIP_OK (0) or an error code, and sets a thread-local message readable with
 *     noa_hip_last_error(). Validation happens before anything is launched.
 *   - Accumulating outputs must be ordinary device memory (hipMalloc); device-wide float atomics are used.
 */
#ifndef NOA_HIP_H
#define NOA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NOA_HIP_ABI_VERSION 1

/* Status codes. */
enum {
    NOA_HIP_OK = 0,
    NOA_HIP_ERROR_INVALID_ARGUMENT = 1,  /* the check_parameters rules (FourierProject.hpp:1095-1185) */
    NOA_HIP_ERROR_UNSUPPORTED = 2,       /* a valid noa call this library does not implement */
    NOA_HIP_ERROR_RUNTIME = 3            /* a HIP runtime error (launch, allocation, ...) */
};

/* Element types of a view. */
enum {
    NOA_HIP_DTYPE_NONE = 0,
    NOA_HIP_F32 = 1,
    NOA_HIP_F64 = 2,
    NOA_HIP_C32 = 3,      /* Complex<f32> */
    NOA_HIP_C64 = 4,      /* Complex<f64> */
    NOA_HIP_C16 = 5,      /* {f16 real, f16 imag}: storage-only volume format for extraction (FP32 math) */
    NOA_HIP_CBF16 = 6,    /* {bf16 real, bf16 imag}: storage-only volume format for extraction (FP32 math) */
    NOA_HIP_C32_BLOCKS = 7 /* Complex<f32> volume in the 4x2x4-block layout of noa_hip_volume_blocks: accepted as the
                            * volume of batched extractions only; the view's shape is the volume's rfft shape and its
                            * strides those of a contiguous volume (the layout fixes the element offsets) */
};

/* noa::fft::Layout values (src/noa/fft/core/Layout.hpp:39-68): the half-to-half layouts, and the full-to-full ones
 * that the CTF functions also accept. */
enum {
    NOA_HIP_H2H = 0,
    NOA_HIP_HC2H = 4,
    NOA_HIP_H2HC = 8,
    NOA_HIP_HC2HC = 12,
    NOA_HIP_F2F = 3,
    NOA_HIP_FC2F = 7,
    NOA_HIP_F2FC = 11,
    NOA_HIP_FC2FC = 15
};

/* noa::xform::Interp::Method values (src/noa/xform/core/Interp.hpp:17-41). */
enum {
    NOA_HIP_INTERP_NEAREST = 0,
    NOA_HIP_INTERP_LINEAR = 1,
    NOA_HIP_INTERP_CUBIC = 2,
    NOA_HIP_INTERP_CUBIC_BSPLINE = 3,
    NOA_HIP_INTERP_LANCZOS4 = 4,
    NOA_HIP_INTERP_LANCZOS6 = 5,
    NOA_HIP_INTERP_LANCZOS8 = 6,
    NOA_HIP_INTERP_NEAREST_FAST = 100,
    NOA_HIP_INTERP_LINEAR_FAST = 101,
    NOA_HIP_INTERP_CUBIC_FAST = 102,
    NOA_HIP_INTERP_CUBIC_BSPLINE_FAST = 103,
    NOA_HIP_INTERP_LANCZOS4_FAST = 104,
    NOA_HIP_INTERP_LANCZOS6_FAST = 105,
    NOA_HIP_INTERP_LANCZOS8_FAST = 106
};

/* noa::fft::Norm and noa::fft::Sign (src/noa/fft/core/Transform.hpp:6-20). */
enum {
    NOA_HIP_NORM_FORWARD = 0,
    NOA_HIP_NORM_ORTHO = 1,
    NOA_HIP_NORM_BACKWARD = 2,
    NOA_HIP_NORM_NONE = 3
};
enum {
    NOA_HIP_SIGN_FORWARD = -1,
    NOA_HIP_SIGN_BACKWARD = 1
};

/* Transform kinds (the Rotate and Scale template arguments, FourierProject.hpp:1086-1091). */
enum {
    NOA_HIP_XFORM_NONE = 0,       /* Empty */
    NOA_HIP_MAT33_F32 = 1,
    NOA_HIP_MAT33_F64 = 2,
    NOA_HIP_QUAT_F32 = 3,
    NOA_HIP_QUAT_F64 = 4,
    NOA_HIP_MAT22_F32 = 5,
    NOA_HIP_MAT22_F64 = 6
};

/* A strided 4D array, an Empty (data == NULL and is_value == 0), or a constant value (is_value == 1),
 * which noa accepts for the inputs of rasterisation and insertion (FourierProject.hpp:1063-1078). */
typedef struct noa_hip_view {
    void* data;
    int64_t shape[4];     /* BDHW rfft shape */
    int64_t strides[4];   /* element strides */
    int32_t dtype;        /* NOA_HIP_F32, ... */
    int32_t is_value;     /* 1: every element equals `value` (data is ignored) */
    double value[2];      /* {real, imag} of the constant value */
} noa_hip_view;

/* One transform broadcast over the batch (batched == 0; `data` is a HOST pointer to one element, copied at
 * the call), or one transform per slice (batched == 1; `data` is a DEVICE pointer to `count` contiguous
 * elements, count == number of slices), as noa's BatchedParameter (src/noa/runtime/core/Utils.hpp:391-413). */
typedef struct noa_hip_transform {
    const void* data;
    int32_t kind;         /* NOA_HIP_MAT33_F32, ... or NOA_HIP_XFORM_NONE */
    int32_t batched;
    int64_t count;
} noa_hip_transform;

/* Arguments of one projection call (one descriptor for the three operators). */
typedef struct noa_hip_project_args {
    noa_hip_view input;            /* slices (rasterize/insert) or volume (extract) */
    noa_hip_view input_weight;     /* optional real weights of the input */
    noa_hip_view output;           /* volume (rasterize/insert) or slices (extract) */
    noa_hip_view output_weight;    /* optional real weights of the output */
    int64_t input_shape[4];        /* logical BDHW shape of the input */
    int64_t output_shape[4];       /* logical BDHW shape of the output */
    int64_t target_shape[4];       /* logical BDHW target shape; all zeros = the volume shape */
    int32_t remap;                 /* NOA_HIP_H2H, NOA_HIP_HC2H, NOA_HIP_H2HC, NOA_HIP_HC2HC */
    int32_t interp;                /* NOA_HIP_INTERP_*; ignored by rasterisation */
    noa_hip_transform rotation;    /* fwd (rasterize, extract) or inv (insert) rotations; required */
    noa_hip_transform scaling;     /* inv (rasterize, extract) or fwd (insert) Mat22 scaling; optional */
    double fftfreq_cutoff;         /* default 0.5 */
    double fftfreq_sinc;           /* windowed sinc (insert: along the slice normal; extract: along w); -1 = default */
    double fftfreq_blackman;       /* -1 = default */
    double ews_radius[2];          /* HW Ewald-sphere radius in 1/pixels; {0,0} = projections */
} noa_hip_project_args;

/* Per-batch 2D shifts of noa::signal::phase_shift_2d (src/noa/signal/PhaseShift.hpp:215-289): (y, x) in pixels,
 * Vec<f32, 2> or Vec<f64, 2> (the coordinate precision). One shift broadcast over the batch (batched == 0; `data`
 * is a HOST pointer to one pair, copied at the call) or one per batch (batched == 1; `data` is a DEVICE pointer to
 * `count` contiguous pairs). Frequencies with |f| > fftfreq_cutoff are not shifted (noa's default cutoff: 1). */
typedef struct noa_hip_shift2 {
    const void* data;
    int32_t dtype;            /* NOA_HIP_F32 or NOA_HIP_F64 */
    int32_t batched;
    int64_t count;
    double fftfreq_cutoff;
} noa_hip_shift2;

/* Arguments of noa::xform::insert_and_extract_central_slices_3d (src/noa/xform/FourierProject.hpp:1916-2049,
 * options :1916-1946): the input slices are inserted into a virtual volume, which is sampled at the frequencies
 * of the output slices without ever being stored. */
typedef struct noa_hip_insert_extract_args {
    noa_hip_view input;                 /* (B_in, 1, H, W/2+1) slices to insert, or a constant value */
    noa_hip_view input_weight;          /* optional real weights of the input slices (array or value) */
    noa_hip_view output;                /* (B_out, 1, H', W'/2+1) extracted slices */
    noa_hip_view output_weight;         /* optional real weights of the extracted slices */
    int64_t input_shape[4];             /* logical BDHW of the input slices */
    int64_t output_shape[4];            /* logical BDHW of the output slices */
    int32_t remap;                      /* input layout -> output layout: H2H, HC2H, H2HC or HC2HC */
    int32_t interp;                     /* NOA_HIP_INTERP_*: interpolation of the input slices */
    noa_hip_transform input_scaling;    /* fwd Mat22 scaling of the inserted slices; optional */
    noa_hip_transform input_rotation;   /* inv rotations of the inserted slices (B_in); required */
    noa_hip_transform output_scaling;   /* inv Mat22 scaling of the extracted slices; optional */
    noa_hip_transform output_rotation;  /* fwd rotations of the extracted slices (B_out); required */
    double input_fftfreq_sinc;          /* input_windowed_sinc, along the normal of each input slice; -1 = default */
    double input_fftfreq_blackman;
    double w_fftfreq_sinc;              /* w_windowed_sinc, along the depth of the virtual volume; -1 = default */
    double w_fftfreq_blackman;
    double fftfreq_cutoff;              /* cutoff of the virtual volume; default 0.5 */
    double ews_radius[2];               /* HW Ewald-sphere radius in 1/pixels; {0,0} = projections */
    int32_t add_to_output;              /* add to the output slices and weights instead of overwriting them */
    int32_t correct_weights;            /* divide the values by max(1, |weight|) (3D only, not with add_to_output) */
} noa_hip_insert_extract_args;

/* noa::signal::CTFAnisotropic<f32> and <f64> as stored by noa (src/noa/signal/core/CTF.hpp:564-575, member order;
 * Vec<f32,2> is 8-byte aligned, Vec<f64,2> 16-byte aligned): 56 and 112 bytes. The members are noa's internal
 * values (defocus in angstroms with noa's sign, cs in angstroms, voltage in volts, bfactor / 4 and the derived k1,
 * k2, k3); a binding passes noa's objects as they are. */
typedef struct noa_hip_ctf_anisotropic_f32 {
    float pixel_size[2];
    float defocus_value, defocus_astigmatism, defocus_angle;
    float phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
} noa_hip_ctf_anisotropic_f32;
typedef struct noa_hip_ctf_anisotropic_f64 {
    double pixel_size[2];
    double defocus_value, defocus_astigmatism, defocus_angle;
    double phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
} noa_hip_ctf_anisotropic_f64;

/* noa::signal::CTFIsotropic<f32> and <f64> as stored by noa (src/noa/signal/core/CTF.hpp:259-271, member order):
 * 44 and 88 bytes, the same internal values as the anisotropic objects with one pixel size and one defocus. */
typedef struct noa_hip_ctf_isotropic_f32 {
    float pixel_size, defocus, phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
} noa_hip_ctf_isotropic_f32;
typedef struct noa_hip_ctf_isotropic_f64 {
    double pixel_size, defocus, phase_shift, cs, voltage, amplitude, bfactor_forth, scale, k1, k2, k3;
} noa_hip_ctf_isotropic_f64;

/* CTFs of a ctf call: one CTF broadcast over the batch (batched == 0; `data` is a HOST pointer to one object, copied
 * at the call) or one per batch (batched == 1; `data` is a DEVICE pointer to `count` objects). The objects are
 * noa_hip_ctf_anisotropic_* for the ctf_anisotropic functions and noa_hip_ctf_isotropic_* for ctf_isotropic. */
typedef struct noa_hip_ctf {
    const void* data;
    int32_t dtype;          /* NOA_HIP_F32 (the f32 objects) or NOA_HIP_F64 */
    int32_t batched;
    int64_t count;
} noa_hip_ctf;

/* CTFOptions (src/noa/signal/CTF.hpp:196-209): fftfreq_range {start, stop, endpoint}, ctf_abs, ctf_squared. */
typedef struct noa_hip_ctf_options {
    double fftfreq_start;   /* must be 0 for 2D spectra */
    double fftfreq_stop;    /* <= 0: the highest frequency of each dimension (default -1) */
    int32_t endpoint;       /* default 1 */
    int32_t ctf_abs;
    int32_t ctf_squared;
} noa_hip_ctf_options;

/* Fill `args` with noa's defaults (cutoff 0.5, sinc/blackman -1, LINEAR, H2H, no scaling). */
void noa_hip_project_args_init(noa_hip_project_args* args);

/* Fill `args` with the defaults of InsertAndExtractCentralSlicesOptions (FourierProject.hpp:1916-1946). */
void noa_hip_insert_extract_args_init(noa_hip_insert_extract_args* args);

/* Library version and ABI version. */
int noa_hip_abi_version(void);

/* Message of the last error raised on the calling thread ("" if none). */
const char* noa_hip_last_error(void);

/* Replaces the GPU branch of noa::xform::rasterize_central_slices_3d (FourierProject.hpp:1653-1692). */
int noa_hip_rasterize_central_slices_3d(const noa_hip_project_args* args, void* stream);

/* Replaces the GPU branch of noa::xform::insert_central_slices_3d (FourierProject.hpp:1770-1815). */
int noa_hip_insert_central_slices_3d(const noa_hip_project_args* args, void* stream);

/* Size of the device scratch array of noa_hip_insert_central_slices_3d_noalloc (one per GPU, held until the process
 * ends). */
#define NOA_HIP_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in noa: noa_hip_insert_central_slices_3d for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation). Same arguments, rules and results
 * (bit-identical). The work plan, the per-slice records and the partial sums of split sub-bricks are carved from a
 * static device array of NOA_HIP_NOALLOC_SCRATCH_BYTES per GPU instead of being allocated; a call whose work needs
 * more returns NOA_HIP_ERROR_UNSUPPORTED with the size in noa_hip_last_error() (at 256^3, up to about 5,000 slices
 * fit). Every call on a GPU shares that GPU's array: calls that overlap on one GPU (different streams, or
 * threads) race and must be ordered by the caller. A geometry's first call builds its plan on the host. */
int noa_hip_insert_central_slices_3d_noalloc(const noa_hip_project_args* args, void* stream);

/* Replaces the GPU branch of noa::xform::extract_central_slices_3d (FourierProject.hpp:1869-1914). */
int noa_hip_extract_central_slices_3d(const noa_hip_project_args* args, void* stream);

/* ---- Batched calls (no counterpart in noa, whose only asynchrony is Stream::enqueue,
 * runtime/cuda/Stream.hpp:205-239) --------------------------------------------------------------------------------
 * Several extraction (or insertion) calls, each with its own arrays, box size, layout, slice count and cutoff, run
 * as ONE persistent launch whose grid does not depend on the batch, so a launch can be captured once in a HIP graph
 * and replayed for other batches. Each call is described by the same noa_hip_project_args as a single call and
 * checked by the same rules; results equal those of the calls made one by one. A batch is made in two steps:
 *   1. *_batch_prepare (host only, nothing launched): writes the batch's device image (header, per-call parameters,
 *      work items) into host memory `image` of `capacity` bytes and its size into *bytes (with image == NULL or a
 *      small capacity: only the size, and NOA_HIP_ERROR_INVALID_ARGUMENT).
 *   2. the caller copies the image to device memory (aligned to 256 bytes), then *_batch_launch enqueues the
 *      kernel(s) on `stream`; the image must stay unchanged until they have run.
 * Batched extraction: complex FP32 volumes and slices (unit x strides), LINEAR, no weights, scaling, Ewald sphere,
 * w window or shifts. Batched insertion: complex FP32 slice arrays (+ optional FP32 weight arrays) into complex FP32
 * volumes (+ optional FP32 weights), LINEAR, no scaling or Ewald sphere; calls that write the same volume race.
 * Grid sizes and pixels per lane come from the launch tuning (NOA_HIP_OP_EXTRACT_BATCH, NOA_HIP_OP_INSERT_BATCH). */
int noa_hip_extract_batch_prepare(const noa_hip_project_args* calls, int32_t n_calls, void* image, int64_t capacity,
                                  int64_t* bytes);
int noa_hip_extract_batch_launch(const void* device_image, void* stream);
int noa_hip_insert_batch_prepare(const noa_hip_project_args* calls, int32_t n_calls, void* image, int64_t capacity,
                                 int64_t* bytes);
int noa_hip_insert_batch_launch(const void* device_image, void* stream);

/* ---- Prepared volumes (no counterpart in noa's projection API; like noa's Texture, xform/Texture.hpp:62-300, a
 * copy made once and read by many calls). For callers that extract from the same resident volume many times, such
 * as a serving loop: the 4x2x4-block copy that the single-call extraction makes on every call, made once, then
 * passed to noa_hip_extract_batch_prepare as a NOA_HIP_C32_BLOCKS view. ---------------------------------------- */
/* Bytes of the block copy of a c32 rfft volume of rfft shape (1, D, H, W/2+1). */
int64_t noa_hip_volume_blocks_bytes(const int64_t rfft_shape[4]);
/* Writes the block copy of `volume` (c32, rfft shape (1, D, H, W/2+1), unit x stride) into `blocks` (device memory
 * of at least noa_hip_volume_blocks_bytes bytes), asynchronously on `stream`. */
int noa_hip_volume_blocks(const noa_hip_view* volume, void* blocks, int64_t capacity, void* stream);

/* ---- Launch tuning (replaces noa's compile-time launch options: IwiseOptions, runtime/Iwise.hpp:14-60, and the
 * fixed 32x16 block reshaping of the CUDA iwise kernels, runtime/cuda/Iwise.cuh:94-175) ---------------------------
 * A table of launch parameters keyed by (operator, box, layout, interpolation, volume type), filled by an autotuner
 * (noa_amd.autotune) or by hand. A key field of -1 (box: 0) matches anything; the entry with the most exact matches
 * wins, the later one on a tie. Values not set (0) take the defaults.
 *   NOA_HIP_OP_EXTRACT:       values = {volume layout: 0 auto, 1 row-major gather, 2 block-copy gather, 3 LDS tiles;
 *                                       expected crossing slices per tile work item (default 128); -; -}
 *   NOA_HIP_OP_INSERT:        values = {kernel: 0 auto (slab where it applies), 1 one-wave, 2 slab; expected
 *                                       slices per work item (default: at most 512 for the slab kernel, 256 for the one-wave kernel,
 *                                       fewer for calls whose expected work fills fewer than 4,096 items); x taps of a slab row: 0 auto (one load),
 *                                       1 one load, 2 two loads; item order: 0 by work, 1 by first slice, 2 XCD-aware,
 *                                       3 pose-adaptive (passing slices counted on the device first; the call then
 *                                       waits for the count, so it is not for stream capture)}
 *   NOA_HIP_OP_EXTRACT_BATCH: values = {pixels per lane, 1-4 (default 3); workgroups (default 2048); -; -}
 *   NOA_HIP_OP_INSERT_BATCH:  values = {workgroups (default 3072); -; -; -}
 * Environment variables (NOA_HIP_EXTRACT_LAYOUT, NOA_HIP_EXTRACT_PAIRS, NOA_HIP_INSERT_KERNEL, NOA_HIP_INSERT_LIST,
 * NOA_HIP_INSERT_PAIRED, NOA_HIP_INSERT_ORDER) override the table; NOA_HIP_INSERT_SPAN sets the most slices of one
 * insertion work item (default 1024, 0: no bound; not applied by the allocation-free entry point or item order 3). */
enum {
    NOA_HIP_OP_EXTRACT = 1,
    NOA_HIP_OP_INSERT = 2,
    NOA_HIP_OP_EXTRACT_BATCH = 3,
    NOA_HIP_OP_INSERT_BATCH = 4
};
typedef struct noa_hip_tuning {
    int32_t op;          /* NOA_HIP_OP_* */
    int32_t box;         /* logical size of the volume's first axis; 0 = any */
    int32_t remap;       /* NOA_HIP_H2H, ...; -1 = any */
    int32_t interp;      /* NOA_HIP_INTERP_*; -1 = any */
    int32_t dtype;       /* value type of the volume; -1 = any */
    int32_t values[4];
} noa_hip_tuning;
/* Replaces the table (count == 0 clears it). */
int noa_hip_set_tuning(const noa_hip_tuning* entries, int32_t count);
/* The values the table gives a key (zeros where nothing matches); returns 1 if an entry matched, else 0. */
int noa_hip_get_tuning(int32_t op, int32_t box, int32_t remap, int32_t interp, int32_t dtype, int32_t values[4]);

/* Replaces the GPU branch of noa::xform::insert_and_extract_central_slices_3d (FourierProject.hpp:1916-2049,
 * launcher :1489-1598, operator FourierInsertExtract :727-1006). Without a w window (w_fftfreq_blackman giving a
 * window of one sample) the output slices are OVERWRITTEN (0 beyond the cutoff) or, with add_to_output, added to;
 * with a w window they are zeroed first unless add_to_output, then the windowed sum over w is added. */
int noa_hip_insert_and_extract_central_slices_3d(const noa_hip_insert_extract_args* args, void* stream);

/* Replaces the GPU branch of noa::signal::ctf_anisotropic<REMAP> (src/noa/signal/CTF.hpp:430-500, operator :14-125)
 * for (B, 1, H, W) spectra, half (output (B, 1, H, W/2+1)) or full layouts: output[b, y, x] = input[b, remap(y, x)] *
 * ctf_b(f) (|.|^2 for a complex input and a real output), or the CTF itself when `input` is Empty (data == NULL).
 * f is the frequency times the step of the fftfreq range; the CTF is squared (ctf_squared) or its absolute value
 * taken (ctf_abs). A half-layout output must be row-major along H, W; a full-layout output may have any strides (H
 * and W are swapped first when W is not the innermost, as noa does). The input is broadcast to the output's shape.
 * The output may be the input when the layout does not change. */
int noa_hip_ctf_anisotropic_2d(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                               int32_t remap, const noa_hip_ctf* ctf, const noa_hip_ctf_options* options,
                               void* stream);

/* Replaces the GPU branch of noa::signal::ctf_isotropic<REMAP> (src/noa/signal/CTF.hpp:216-421, operator :14-125):
 * 1D, 2D or 3D spectra (by the dimensionality of the logical BDHW `shape`), half (output shape with W/2+1) or full
 * layouts, CTFIsotropic objects. output = input[remap] * ctf_b(|f|) (1D: ctf_b(f + start), the only case that
 * allows a non-zero start), |.|^2 for a complex input and a real output, or the CTF itself when `input` is Empty.
 * The dimensions are put in the output's rightmost order first, as noa does; a half-layout output whose W is then
 * not innermost is rejected (noa's loop would leave the array). The input is broadcast to the output's shape. */
int noa_hip_ctf_isotropic(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                          int32_t remap, const noa_hip_ctf* ctf, const noa_hip_ctf_options* options, void* stream);

/* Both CTF steps before a CTF-weighted insertion in one pass: output = input * ctf and output_weight = ctf^2, the
 * results of noa's ctf_anisotropic<REMAP>(input, output, ...) followed by ctf_anisotropic<REMAP>({}, output_weight,
 * ..., {.ctf_squared = true}) with the same CTFs and range (ctf_abs and ctf_squared of `options` are ignored).
 * Extension of this library: the input is read once and the CTF evaluated once per element. */
int noa_hip_ctf_anisotropic_2d_weighted(const noa_hip_view* input, const noa_hip_view* output,
                                        const noa_hip_view* output_weight, const int64_t shape[4], int32_t remap,
                                        const noa_hip_ctf* ctf, const noa_hip_ctf_options* options, void* stream);

/* Replaces the GPU branch of noa::signal::phase_shift_2d<REMAP> (src/noa/signal/PhaseShift.hpp:215-289, operator
 * :78-139) for the rfft layouts: output[b, remap(y, x)] = input[b, y, x] * exp(-2 pi i shift_b . f(y, x)), f the
 * normalised frequency. `input` may be an Empty view (data == NULL): the phase shifts themselves are written.
 * `shape` is the logical BDHW shape with D == 1; input and output are c32 or c64 (B, 1, H, W/2+1) arrays, and may be
 * the same array when the layout does not change. A single zero shift copies (or remaps), a single shift of half
 * the shape uses the exact +-1 pattern (:250-267), as in noa. */
int noa_hip_phase_shift_2d(const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                           int32_t remap, const noa_hip_shift2* shifts, void* stream);

/* Extraction followed by noa::signal::phase_shift_2d of the extracted slices in their own layout (hc2hc when the
 * slices are centred, h2h otherwise), fused: the phase shift is applied before the slices are stored, so the
 * slices are written once. Same arguments and rules as noa_hip_extract_central_slices_3d, plus the shifts (one
 * per slice or one broadcast); extension of this library, no single noa function does both. */
int noa_hip_extract_central_slices_3d_shifted(const noa_hip_project_args* args, const noa_hip_shift2* shifts,
                                              void* stream);

/* Replaces the GPU branch of noa::fft::remap (src/noa/fft/Remap.hpp:205-340) for the half-to-half layouts (h2h,
 * h2hc, hc2h, hc2hc) of (B, D, H, W/2+1) rfft arrays: output[b, remap(z, y), x] = input[b, z, y, x], the D and H
 * indices fftshifted or ifftshifted. Input and output have the same element type; they must not overlap when the
 * layout changes (a layout-preserving remap is a copy and may be in place). */
int noa_hip_remap(int32_t remap, const noa_hip_view* input, const noa_hip_view* output, const int64_t shape[4],
                  void* stream);

/* Replaces the GPU branch of noa::xform::fourier_interpolation_correction (src/noa/xform/FourierProject.hpp:
 * 2062-2087, operator :1008-1061): real-space correction of linear interpolation, output = input / sinc^2(pi r)
 * (post_correction != 0) or input * sinc^2(pi r), r the distance to the box centre in fractions of the box. Real
 * f32/f64 (B, D, H, W) arrays; the input may have a batch of 1 (broadcast) and may be the output. interp must be
 * LINEAR or LINEAR_FAST, as in noa. */
int noa_hip_fourier_interpolation_correction(const noa_hip_view* input, const noa_hip_view* output, int32_t interp,
                                             int32_t post_correction, void* stream);

/* noa::fft::FFTOptions (src/noa/fft/Transform.hpp:39-88); noa_hip_fft_options_init sets noa's defaults
 * (norm FORWARD, cache_plan 1, plan_only 0, record_and_share_workspace 0). */
typedef struct noa_hip_fft_options {
    int32_t norm;
    int32_t cache_plan;
    int32_t plan_only;
    int32_t record_and_share_workspace;
} noa_hip_fft_options;

void noa_hip_fft_options_init(noa_hip_fft_options* options);

/* Replace the GPU branches of noa::fft::r2c, c2r and c2c (src/noa/fft/Transform.hpp:160-349; noa's CUDA path
 * src/noa/fft/cuda/Transform.hpp, Plan.cpp) with rocFFT. (B, D, H, W) views with element strides (non-negative);
 * r2c: real f32/f64 input of the logical shape into a complex output of shape (B, D, H, W/2+1) ("h" layout);
 * c2r: the reverse (for 2D and 3D transforms the complex input is overwritten, as in noa); c2c: complex arrays of
 * equal shapes, sign NOA_HIP_SIGN_FORWARD or _BACKWARD. Equal data pointers run in place (r2c/c2r then need noa's
 * padded layout: the complex strides are half the real ones and the real rows hold 2 * (W/2+1) elements). The
 * normalisation is noa's (1/N forward by default, N = D*H*W), applied by the transform itself. options may be NULL
 * (noa's defaults). Plans are cached per host thread and device as in noa (4 plans by default). */
int noa_hip_fft_r2c(const noa_hip_view* input, const noa_hip_view* output, const noa_hip_fft_options* options,
                    void* stream);
int noa_hip_fft_c2r(const noa_hip_view* input, const noa_hip_view* output, const noa_hip_fft_options* options,
                    void* stream);
int noa_hip_fft_c2c(const noa_hip_view* input, const noa_hip_view* output, int32_t sign,
                    const noa_hip_fft_options* options, void* stream);

/* The plan cache of the calling host thread on `device` (noa::fft::clear_cache, set_cache_limit, cache_limit,
 * cache_size, workspace_left_to_allocate, set_workspace; src/noa/fft/Transform.hpp:90-142). clear_cache returns
 * the number of plans released, set_cache_limit the number evicted, set_workspace the number of plans that took the
 * caller's device buffer (which must outlive them), 0 when it is smaller than workspace_left_to_allocate. */
int32_t noa_hip_fft_clear_cache(int32_t device);
int32_t noa_hip_fft_set_cache_limit(int32_t count, int32_t device);
int32_t noa_hip_fft_cache_limit(int32_t device);
int32_t noa_hip_fft_cache_size(int32_t device);
int64_t noa_hip_fft_workspace_left_to_allocate(int32_t device);
int32_t noa_hip_fft_set_workspace(int32_t device, void* buffer, int64_t bytes);

/* noa::io::DataType values (src/noa/io/Encoding.hpp:19-26): the on-disk encodings of MRC files. */
enum {
    NOA_HIP_IO_UNKNOWN = 0,
    NOA_HIP_IO_I8 = 1, NOA_HIP_IO_U8 = 2, NOA_HIP_IO_I16 = 3, NOA_HIP_IO_U16 = 4, NOA_HIP_IO_I32 = 5,
    NOA_HIP_IO_U32 = 6, NOA_HIP_IO_I64 = 7, NOA_HIP_IO_U64 = 8,
    NOA_HIP_IO_F16 = 9, NOA_HIP_IO_F32 = 10, NOA_HIP_IO_F64 = 11,
    NOA_HIP_IO_C16 = 12, NOA_HIP_IO_C32 = 13, NOA_HIP_IO_C64 = 14,
    NOA_HIP_IO_U4 = 15, NOA_HIP_IO_CI16 = 16
};

/* What noa's MRC encoder reads from and writes to the 1024-byte header (src/noa/io/ImageFile.cpp:14-166, :169-296):
 * BDHW shape (a 2D image, a 3D volume (space group 1), a stack of 2D images (group 0) or of volumes (groups
 * 401-630)), DHW spacing in Angstrom per pixel (0 = unset), the data type of the stored values (mode 0 -> I8, or U8
 * with the IMOD flag; 1 I16, 2 F32, 3 CI16, 4 C32, 6 U16, 12 F16, 101 U4), the statistics when set, and where the
 * data starts (1024 + the extended header). */
typedef struct noa_hip_mrc_header {
    int64_t shape[4];
    double spacing[3];
    int32_t dtype;            /* NOA_HIP_IO_* */
    int32_t has_stats;
    double stats[4];          /* min, max, mean, stddev */
    int64_t data_offset;      /* bytes */
    int32_t endian_swapped;   /* the file is big-endian (on this little-endian host) */
    int32_t reserved;
} noa_hip_mrc_header;

/* Parses the header of an MRC file (noa::io::ImageFile(path, {.read = true}) with the MRC encoder). */
int noa_hip_mrc_read_header(const char* path, noa_hip_mrc_header* header);

/* noa::io::read_image of an MRC file into device memory (src/noa/io/IO.hpp:34-57): `output` is a f32, f64, c32 or
 * c64 device view of the file's BDHW shape (any strides). The file is read in chunks into pinned host buffers and
 * copied to the device while the next chunk is read; the values are decoded on the GPU with noa's rules (byte swap
 * of big-endian files, then clamp_cast to the output type when clamp != 0, src/noa/io/Encoding.cpp:85-99,
 * base/ClampCast.hpp). Complex files need a complex output; real files into a complex output give {value, 0}.
 * Returns when the data is in `output` (the stream is synchronised). */
int noa_hip_mrc_read(const char* path, const noa_hip_view* output, int32_t clamp, void* stream);

/* noa::io::write_image into a new MRC file (src/noa/io/IO.hpp:119-163): `input` is a f32, f64, c32 or c64 device
 * view; dtype is the requested NOA_HIP_IO_* encoding (UNKNOWN: F32 for real, C32 for complex input), replaced by
 * the closest one MRC supports as noa does (ImageFile.hpp:606-628); spacing is DHW (NULL: unset); stats is NULL or
 * {min, max, mean, stddev}. Values are encoded on the GPU (clamp_cast to the file type when clamp != 0) and written
 * in chunks through pinned host buffers. Returns when the file is written and closed. */
int noa_hip_mrc_write(const char* path, const noa_hip_view* input, const double* spacing, int32_t dtype,
                      const double* stats, int32_t clamp, void* stream);

/* Replaces the GPU branch of noa::xform::cubic_bspline_prefilter (src/noa/xform/CubicBSplinePrefilter.hpp:24-63,
 * algorithm core/CubicBSplinePrefilter.hpp:22-107): the recursive cubic B-spline prefilter along x, then y, then z
 * of the dimensions reordered to the output's rightmost order, so that CUBIC_BSPLINE interpolation of the output
 * passes through the input samples. (B, D, H, W) views (shape and element strides) of f32, f64, c32 or c64; the
 * input is broadcast into the output's shape (dimensions of size 1) and may be the output (in place, same data
 * pointer); any other overlap is rejected. */
int noa_hip_cubic_bspline_prefilter(const noa_hip_view* input, const noa_hip_view* output, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NOA_HIP_H */
