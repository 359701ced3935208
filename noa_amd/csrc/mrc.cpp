// MRC files between disk and device memory: noa's MRC encoder (src/noa/io/ImageFile.hpp:558-747, header codec
// src/noa/io/ImageFile.cpp:14-296) with noa::io::read_image / write_image (src/noa/io/IO.hpp:34-163).
//
// The header is parsed and written on the host, field by field as noa does. The data moves in 32 MiB chunks
// through two pinned host buffers: while the GPU copies and decodes chunk k (mrc_codec.hip), the host reads chunk
// k + 1 from the file; writing runs the other way (encode on the GPU, copy back, then fwrite while the next chunk
// is encoded). The values never take a CPU conversion pass.
//
// One deliberate difference: for big-endian files noa reverses the first 24 header bytes before swapping each
// 4-byte word (ImageFile.cpp:34-39), which scrambles nx, ny, nz, mode, nxstart and nystart; here every word is
// swapped in place, so big-endian files read correctly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cstdarg>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/noa_hip.h"
#include "kernels.hpp"

namespace noa_hip {
namespace {

using u32 = uint32_t;

constexpr i64 HEADER_SIZE = 1024;
constexpr i64 CHUNK_BYTES = 32 << 20;

struct Fail { int code; std::string msg; };

[[noreturn]] void fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Fail{code, buf};
}

// DataType::n_bytes (Encoding.hpp:85-113).
i64 n_bytes(int dtype, i64 n) {
    switch (dtype) {
        case NOA_HIP_IO_U4: return n / 2;
        case NOA_HIP_IO_I8: case NOA_HIP_IO_U8: return n;
        case NOA_HIP_IO_I16: case NOA_HIP_IO_U16: case NOA_HIP_IO_F16: return 2 * n;
        case NOA_HIP_IO_I32: case NOA_HIP_IO_U32: case NOA_HIP_IO_F32: case NOA_HIP_IO_CI16: case NOA_HIP_IO_C16: return 4 * n;
        case NOA_HIP_IO_I64: case NOA_HIP_IO_U64: case NOA_HIP_IO_F64: case NOA_HIP_IO_C32: return 8 * n;
        case NOA_HIP_IO_C64: return 16 * n;
        default: return 0;
    }
}

bool is_complex_io(int dtype) {
    return dtype == NOA_HIP_IO_C16 || dtype == NOA_HIP_IO_C32 || dtype == NOA_HIP_IO_C64 || dtype == NOA_HIP_IO_CI16;
}

// ImageFileEncoderMrc::closest_supported_dtype (ImageFile.hpp:606-628).
int closest_supported(int dtype) {
    switch (dtype) {
        case NOA_HIP_IO_I8: case NOA_HIP_IO_U8: case NOA_HIP_IO_I16: case NOA_HIP_IO_U16: case NOA_HIP_IO_F16:
        case NOA_HIP_IO_U4: case NOA_HIP_IO_CI16:
            return dtype;
        case NOA_HIP_IO_I32: case NOA_HIP_IO_U32: case NOA_HIP_IO_I64: case NOA_HIP_IO_U64: case NOA_HIP_IO_F32:
        case NOA_HIP_IO_F64:
            return NOA_HIP_IO_F32;
        case NOA_HIP_IO_C16: case NOA_HIP_IO_C32: case NOA_HIP_IO_C64:
            return NOA_HIP_IO_C32;
        default:
            return NOA_HIP_IO_UNKNOWN;
    }
}

u32 bswap32(u32 v) { return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24); }

template<class T> T get(const unsigned char* b, int off) { T v; std::memcpy(&v, b + off, sizeof(T)); return v; }
template<class T> void put(unsigned char* b, int off, T v) { std::memcpy(b + off, &v, sizeof(T)); }

struct File {
    std::FILE* f{};
    ~File() { if (f) std::fclose(f); }
};

// ImageFileEncoderMrc::read_header (ImageFile.cpp:14-166).
noa_hip_mrc_header parse_header(std::FILE* f) {
    unsigned char b[1024];
    if (std::fseek(f, 0, SEEK_END) != 0) fail(NOA_HIP_ERROR_RUNTIME, "Failed to seek. %s", std::strerror(errno));
    const long size = std::ftell(f);
    if (std::fseek(f, 0, SEEK_SET) != 0 || size < 0) fail(NOA_HIP_ERROR_RUNTIME, "Failed to seek. %s", std::strerror(errno));
    if (size < 1024 || std::fread(b, 1, 1024, f) != 1024)
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Failed to read the header (1024 bytes)");

    noa_hip_mrc_header h{};
    const unsigned char* st = b + 212;
    if ((st[0] == 68 && st[1] == 65 && st[2] == 0 && st[3] == 0) || (st[0] == 68 && st[1] == 68 && st[2] == 0 && st[3] == 0)) {
        h.endian_swapped = 0;
    } else if (st[0] == 17 && st[1] == 17 && st[2] == 0 && st[3] == 0) {
        h.endian_swapped = 1;
    } else {
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
             "Invalid data. Endianness was not recognized. Should be [68,65,0,0], [68,68,0,0] or [17,17,0,0], "
             "got [%d,%d,%d,%d]", st[0], st[1], st[2], st[3]);
    }
    if (h.endian_swapped) {
        // Words 0-23 (nx .. next), imodStamp/imodFlags and rms/nlabl.
        auto swap_words = [&](int off, int n) {
            for (int i = 0; i < n; ++i) put<u32>(b, off + 4 * i, bswap32(get<u32>(b, off + 4 * i)));
        };
        swap_words(0, 24);
        swap_words(152, 2);
        swap_words(216, 2);
    }
    const int32_t nx = get<int32_t>(b, 0), ny = get<int32_t>(b, 4), nz = get<int32_t>(b, 8);
    const int32_t mode = get<int32_t>(b, 12);
    const int32_t mx = get<int32_t>(b, 28), my = get<int32_t>(b, 32), mz = get<int32_t>(b, 36);
    const float cx = get<float>(b, 40), cy = get<float>(b, 44), cz = get<float>(b, 48);
    const int32_t mapc = get<int32_t>(b, 64), mapr = get<int32_t>(b, 68), maps = get<int32_t>(b, 72);
    const float dmin = get<float>(b, 76), dmax = get<float>(b, 80), dmean = get<float>(b, 84);
    const int32_t ispg = get<int32_t>(b, 88);
    const int32_t next = get<int32_t>(b, 92);
    const int32_t imod_stamp = get<int32_t>(b, 152), imod_flags = get<int32_t>(b, 156);
    const float rms = get<float>(b, 216);

    if (nx <= 0 || ny <= 0 || nz <= 0)
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
             "Invalid data. Logical shape should be greater than zero, got nx,ny,nz:(%d,%d,%d)", nx, ny, nz);
    // Shape (:73-117): ndim of (nz, ny, nx).
    const int ndim = nz > 1 ? 3 : (ny > 1 && nx > 1) ? 2 : 1;
    if (ndim <= 2) {
        if (!(mx == nx && my == ny && mz == nz))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                 "1d or 2d data detected. The logical shape should be equal to the grid size. Got nx,ny,nz:(%d,%d,%d), "
                 "mx,my,mz:(%d,%d,%d)", nx, ny, nz, mx, my, mz);
        h.shape[0] = 1; h.shape[1] = 1; h.shape[2] = ny; h.shape[3] = nx;
    } else if (ispg == 0) {
        if (!(nx == mx && ny == my))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                 "2d stack of images detected (ndim=3, group=0). The two innermost dimensions of the logical shape and "
                 "the grid size should be equal. Got nx,ny,nz:(%d,%d,%d), mx,my,mz:(%d,%d,%d)", nx, ny, nz, mx, my, mz);
        h.shape[0] = nz; h.shape[1] = 1; h.shape[2] = ny; h.shape[3] = nx;
    } else if (ispg == 1) {
        if (!(mx == nx && my == ny && mz == nz))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                 "3d volume detected (ndim=3, group=1). The logical shape should be equal to the grid size. "
                 "Got nx,ny,nz:(%d,%d,%d), mx,my,mz:(%d,%d,%d)", nx, ny, nz, mx, my, mz);
        h.shape[0] = 1; h.shape[1] = nz; h.shape[2] = ny; h.shape[3] = nx;
    } else if (ispg >= 401 && ispg <= 630) {
        if (mz <= 0 || nz % mz != 0)
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                 "Stack of 3d volumes detected. The total sections (nz:%d) should be divisible by the number of "
                 "sections per volume (mz:%d)", nz, mz);
        if (!(nx == mx && ny == my))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                 "Stack of 3d volumes detected. The first two dimensions of the shape and the grid size should be equal. "
                 "Got nx,ny,nz:(%d,%d,%d), mx,my,mz:(%d,%d,%d)", nx, ny, nz, mx, my, mz);
        h.shape[0] = nz / mz; h.shape[1] = mz; h.shape[2] = ny; h.shape[3] = nx;
    } else {
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Data shape is not recognized. Got nx,ny,nz:(%d,%d,%d), mx,my,mz:(%d,%d,%d), group:%d",
             nx, ny, nz, mx, my, mz, ispg);
    }
    // Spacing (:119-123): cell size / grid size, DHW, in f32.
    const float sx = cx / static_cast<float>(mx), sy = cy / static_cast<float>(my), sz = cz / static_cast<float>(mz);
    h.spacing[0] = sz; h.spacing[1] = sy; h.spacing[2] = sx;
    if (!(sx >= 0 && sy >= 0 && sz >= 0))
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Invalid data. Pixel size should not be negative, got (%g,%g,%g)", sz, sy, sx);
    if (next < 0)
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Invalid data. Extended header size should be positive, got %d", next);
    h.data_offset = HEADER_SIZE + next;
    // Mode (:126-151).
    switch (mode) {
        case 0: h.dtype = (imod_stamp == 1146047817 && (imod_flags & 1)) ? NOA_HIP_IO_U8 : NOA_HIP_IO_I8; break;
        case 1: h.dtype = NOA_HIP_IO_I16; break;
        case 2: h.dtype = NOA_HIP_IO_F32; break;
        case 3: h.dtype = NOA_HIP_IO_CI16; break;
        case 4: h.dtype = NOA_HIP_IO_C32; break;
        case 6: h.dtype = NOA_HIP_IO_U16; break;
        case 12: h.dtype = NOA_HIP_IO_F16; break;
        case 16: fail(NOA_HIP_ERROR_UNSUPPORTED, "MRC mode 16 is not currently supported");
        case 101:
            if (h.shape[3] % 2 != 0)
                fail(NOA_HIP_ERROR_UNSUPPORTED, "Mode 101 (u4) is only supported for shapes with even width");
            h.dtype = NOA_HIP_IO_U4;
            break;
        default: fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Invalid data. MRC mode not recognized, got %d", mode);
    }
    // Map order (:153-158).
    if (!(mapc == 1 && mapr == 2 && maps == 3)) {
        if (mapc < 1 || mapr < 1 || maps < 1 || mapc > 3 || mapr > 3 || maps > 3 || mapc + mapr + maps != 6)
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Invalid data. Map order should be (1,2,3), got (%d,%d,%d)", mapc, mapr, maps);
        fail(NOA_HIP_ERROR_UNSUPPORTED, "Map order (%d,%d,%d) is not supported. Only (1,2,3) is supported", mapc, mapr, maps);
    }
    // Statistics (:160-166): unset when max == -1 and mean == 2.
    if (!(dmax == -1.f && dmean == 2.f)) {
        h.has_stats = 1;
        h.stats[0] = dmin; h.stats[1] = dmax; h.stats[2] = dmean; h.stats[3] = rms;
    }
    const i64 need = h.data_offset + n_bytes(h.dtype, h.shape[0] * h.shape[1] * h.shape[2] * h.shape[3]);
    if (size < need)
        fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "The file is truncated: %ld bytes for %lld bytes of header and data", size,
             static_cast<long long>(need));
    return h;
}

// ImageFileEncoderMrc::write_header (ImageFile.cpp:169-296), little-endian host.
void write_header(std::FILE* f, const i64 shape[4], const double spacing[3], int dtype, const double* stats) {
    int32_t mode = 0, imod_stamp = 0, imod_flags = 0;
    switch (dtype) {
        case NOA_HIP_IO_U8: mode = 0; imod_stamp = 1146047817; imod_flags &= 1; break;   // as noa: the flag stays 0
        case NOA_HIP_IO_I8: mode = 0; break;
        case NOA_HIP_IO_I16: mode = 1; break;
        case NOA_HIP_IO_F32: mode = 2; break;
        case NOA_HIP_IO_CI16: mode = 3; break;
        case NOA_HIP_IO_C32: mode = 4; break;
        case NOA_HIP_IO_U16: mode = 6; break;
        case NOA_HIP_IO_F16: mode = 12; break;
        case NOA_HIP_IO_U4:
            if (shape[3] % 2 != 0) fail(NOA_HIP_ERROR_UNSUPPORTED, "Mode 101 (u4) is only supported for shapes with even width");
            mode = 101;
            break;
        default: fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "The data type is set to unknown");
    }
    int32_t whd[3], grid[3], group;
    if (shape[0] == 1) {   // 1d, 2d image or 3d volume
        whd[0] = grid[0] = static_cast<int32_t>(shape[3]);
        whd[1] = grid[1] = static_cast<int32_t>(shape[2]);
        whd[2] = grid[2] = static_cast<int32_t>(shape[1]);
        const int ndim = shape[1] > 1 ? 3 : (shape[2] > 1 && shape[3] > 1) ? 2 : 1;
        group = ndim == 3 ? 1 : 0;
    } else if (shape[1] == 1) {   // stack of 2d images
        whd[0] = grid[0] = static_cast<int32_t>(shape[3]);
        whd[1] = grid[1] = static_cast<int32_t>(shape[2]);
        whd[2] = static_cast<int32_t>(shape[0]);
        grid[2] = 1;
        group = 0;
    } else {   // stack of volumes
        whd[0] = grid[0] = static_cast<int32_t>(shape[3]);
        whd[1] = grid[1] = static_cast<int32_t>(shape[2]);
        whd[2] = static_cast<int32_t>(shape[1] * shape[0]);
        grid[2] = static_cast<int32_t>(shape[1]);
        group = 401;
    }
    // cell = grid * spacing (WHD), in f32.
    const float sp[3] = {static_cast<float>(spacing[2]), static_cast<float>(spacing[1]), static_cast<float>(spacing[0])};
    float cell[3];
    for (int i = 0; i < 3; ++i) cell[i] = static_cast<float>(grid[i]) * sp[i];
    float st[4] = {0.f, -1.f, 2.f, -1.f};
    if (stats)
        for (int i = 0; i < 4; ++i) st[i] = static_cast<float>(stats[i]);

    unsigned char b[1024] = {};
    for (int i = 0; i < 3; ++i) put<int32_t>(b, 4 * i, whd[i]);
    put<int32_t>(b, 12, mode);
    for (int i = 0; i < 3; ++i) put<int32_t>(b, 28 + 4 * i, grid[i]);
    for (int i = 0; i < 3; ++i) put<float>(b, 40 + 4 * i, cell[i]);
    for (int i = 0; i < 3; ++i) put<float>(b, 52 + 4 * i, 90.f);
    for (int i = 0; i < 3; ++i) put<int32_t>(b, 64 + 4 * i, i + 1);
    for (int i = 0; i < 3; ++i) put<float>(b, 76 + 4 * i, st[i]);
    put<int32_t>(b, 88, group);
    put<int32_t>(b, 92, 0);
    std::memcpy(b + 104, "SERI", 4);
    put<int32_t>(b, 152, imod_stamp);
    put<int32_t>(b, 156, imod_flags);
    std::memcpy(b + 208, "MAP ", 4);
    b[212] = 68; b[213] = 68;
    put<float>(b, 216, st[3]);
    if (std::fseek(f, 0, SEEK_SET) != 0 || std::fwrite(b, 1, 1024, f) != 1024)
        fail(NOA_HIP_ERROR_RUNTIME, "Failed to write the header (1024 bytes). %s", std::strerror(errno));
}

// Two pinned host buffers and two device buffers of CHUNK_BYTES, and an event per pair.
struct Staging {
    void* host[2]{};
    void* dev[2]{};
    hipEvent_t done[2]{};
    explicit Staging(i64 bytes) {
        for (int j = 0; j < 2; ++j) {
            if (hipHostMalloc(&host[j], static_cast<size_t>(bytes), hipHostMallocDefault) != hipSuccess ||
                hipMalloc(&dev[j], static_cast<size_t>(bytes)) != hipSuccess ||
                hipEventCreateWithFlags(&done[j], hipEventDisableTiming) != hipSuccess) {
                (void)hipGetLastError();
                release();
                fail(NOA_HIP_ERROR_RUNTIME, "mrc: could not allocate the %lld-byte staging buffers", static_cast<long long>(bytes));
            }
        }
    }
    void release() {
        for (int j = 0; j < 2; ++j) {
            if (done[j]) { (void)hipEventSynchronize(done[j]); (void)hipEventDestroy(done[j]); done[j] = nullptr; }
            if (dev[j]) { (void)hipFree(dev[j]); dev[j] = nullptr; }
            if (host[j]) { (void)hipHostFree(host[j]); host[j] = nullptr; }
        }
    }
    ~Staging() { release(); }
};

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(NOA_HIP_ERROR_RUNTIME, "%s: %s", what, hipGetErrorString(e));
}

void check_array(const noa_hip_view& v, const char* name) {
    if (v.data == nullptr || v.is_value) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Empty array detected");
    if (!(v.dtype == NOA_HIP_F32 || v.dtype == NOA_HIP_F64 || v.dtype == NOA_HIP_C32 || v.dtype == NOA_HIP_C64))
        fail(NOA_HIP_ERROR_UNSUPPORTED, "mrc: the %s should be f32, f64, c32 or c64", name);
    for (int i = 0; i < 4; ++i)
        if (v.shape[i] <= 0) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Empty array detected");
}

bool array_is_complex(int dtype) { return dtype == NOA_HIP_C32 || dtype == NOA_HIP_C64; }

MrcChunk chunk_of(const noa_hip_view& v) {
    MrcChunk c;
    c.array = v.data;
    c.array_dtype = v.dtype;
    for (int i = 0; i < 4; ++i) { c.s[i] = v.strides[i]; c.n[i] = v.shape[i]; }
    return c;
}

// Elements per chunk: CHUNK_BYTES of file data, even (4-bit data packs two elements per byte).
i64 chunk_elements(int dtype) {
    const i64 per = n_bytes(dtype, 2);
    return std::max<i64>(2, (CHUNK_BYTES * 2 / per) & ~static_cast<i64>(1));
}

} // namespace

int mrc_read_header(const char* path, noa_hip_mrc_header& h, std::string& err) {
    try {
        if (!path) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "null path");
        File file;
        file.f = std::fopen(path, "rb");
        if (!file.f) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Failed to open %s. %s", path, std::strerror(errno));
        h = parse_header(file.f);
        return NOA_HIP_OK;
    } catch (const Fail& f) {
        err = f.msg;
        return f.code;
    }
}

int mrc_read(const char* path, const noa_hip_view& out, bool clamp, hipStream_t stream, std::string& err) {
    try {
        if (!path) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "null path");
        check_array(out, "output");
        File file;
        file.f = std::fopen(path, "rb");
        if (!file.f) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Failed to open %s. %s", path, std::strerror(errno));
        const noa_hip_mrc_header h = parse_header(file.f);
        for (int i = 0; i < 4; ++i)
            if (out.shape[i] != h.shape[i])
                fail(NOA_HIP_ERROR_INVALID_ARGUMENT,
                     "The output shape (%lld,%lld,%lld,%lld) does not match the file shape (%lld,%lld,%lld,%lld)",
                     (long long)out.shape[0], (long long)out.shape[1], (long long)out.shape[2], (long long)out.shape[3],
                     (long long)h.shape[0], (long long)h.shape[1], (long long)h.shape[2], (long long)h.shape[3]);
        if (is_complex_io(h.dtype) && !array_is_complex(out.dtype))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Cannot decode complex data into a real array");
        if (std::fseek(file.f, static_cast<long>(h.data_offset), SEEK_SET) != 0)
            fail(NOA_HIP_ERROR_RUNTIME, "Failed to seek. %s", std::strerror(errno));

        const i64 total = h.shape[0] * h.shape[1] * h.shape[2] * h.shape[3];
        const i64 per_chunk = chunk_elements(h.dtype);
        Staging sg(n_bytes(h.dtype, per_chunk));
        MrcChunk c = chunk_of(out);
        c.file_dtype = h.dtype;
        c.swap = h.endian_swapped != 0;
        c.clamp = clamp;
        int j = 0;
        for (i64 first = 0; first < total; first += per_chunk, j ^= 1) {
            const i64 count = std::min(per_chunk, total - first);
            const i64 bytes = n_bytes(h.dtype, count);
            check_hip(hipEventSynchronize(sg.done[j]), "mrc: wait for the staging buffer");
            if (std::fread(sg.host[j], 1, static_cast<size_t>(bytes), file.f) != static_cast<size_t>(bytes))
                fail(NOA_HIP_ERROR_RUNTIME, "Failed to read %lld bytes of data", static_cast<long long>(bytes));
            check_hip(hipMemcpyAsync(sg.dev[j], sg.host[j], static_cast<size_t>(bytes), hipMemcpyHostToDevice, stream),
                      "mrc: copy to the device");
            c.raw = sg.dev[j];
            c.first = first;
            c.count = count;
            int st = launch_mrc_decode(c, stream, err);
            if (st != NOA_HIP_OK) return st;
            check_hip(hipEventRecord(sg.done[j], stream), "mrc: record");
        }
        check_hip(hipStreamSynchronize(stream), "mrc: synchronise");
        return NOA_HIP_OK;
    } catch (const Fail& f) {
        err = f.msg;
        return f.code;
    }
}

int mrc_write(const char* path, const noa_hip_view& in, const double* spacing, int dtype, const double* stats,
              bool clamp, hipStream_t stream, std::string& err) {
    try {
        if (!path) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "null path");
        check_array(in, "input");
        if (dtype < NOA_HIP_IO_UNKNOWN || dtype > NOA_HIP_IO_CI16) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "invalid dtype %d", dtype);
        int file_dtype = dtype == NOA_HIP_IO_UNKNOWN
                         ? (array_is_complex(in.dtype) ? NOA_HIP_IO_C32 : NOA_HIP_IO_F32) : closest_supported(dtype);
        if (file_dtype == NOA_HIP_IO_UNKNOWN) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "The data type is set to unknown");
        if (array_is_complex(in.dtype) && !is_complex_io(file_dtype))
            fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Cannot cast a complex value to a non-complex value");
        const double zero[3] = {0, 0, 0};
        i64 shape[4];
        for (int i = 0; i < 4; ++i) shape[i] = in.shape[i];
        for (int i = 0; i < 4; ++i)
            if (shape[i] > 0x7fffffffLL) fail(NOA_HIP_ERROR_UNSUPPORTED, "mrc: dimensions above 2^31 cannot be stored");
        if (shape[0] > 1 && shape[1] > 1 && shape[0] * shape[1] > 0x7fffffffLL)
            fail(NOA_HIP_ERROR_UNSUPPORTED, "mrc: more than 2^31 sections cannot be stored");

        File file;
        file.f = std::fopen(path, "wb");
        if (!file.f) fail(NOA_HIP_ERROR_INVALID_ARGUMENT, "Failed to open %s. %s", path, std::strerror(errno));
        write_header(file.f, shape, spacing ? spacing : zero, file_dtype, stats);

        const i64 total = shape[0] * shape[1] * shape[2] * shape[3];
        const i64 per_chunk = chunk_elements(file_dtype);
        Staging sg(n_bytes(file_dtype, per_chunk));
        MrcChunk c = chunk_of(in);
        c.file_dtype = file_dtype;
        c.clamp = clamp;
        i64 pending_bytes[2] = {0, 0};
        auto flush = [&](int k) {
            if (pending_bytes[k] == 0) return;
            check_hip(hipEventSynchronize(sg.done[k]), "mrc: wait for the encoded chunk");
            if (std::fwrite(sg.host[k], 1, static_cast<size_t>(pending_bytes[k]), file.f) != static_cast<size_t>(pending_bytes[k]))
                fail(NOA_HIP_ERROR_RUNTIME, "Failed to write %lld bytes of data. %s", static_cast<long long>(pending_bytes[k]),
                     std::strerror(errno));
            pending_bytes[k] = 0;
        };
        int j = 0;
        for (i64 first = 0; first < total; first += per_chunk, j ^= 1) {
            flush(j);   // the buffer pair j is free once its previous chunk is on disk
            const i64 count = std::min(per_chunk, total - first);
            const i64 bytes = n_bytes(file_dtype, count);
            c.raw = sg.dev[j];
            c.first = first;
            c.count = count;
            int st = launch_mrc_encode(c, stream, err);
            if (st != NOA_HIP_OK) return st;
            check_hip(hipMemcpyAsync(sg.host[j], sg.dev[j], static_cast<size_t>(bytes), hipMemcpyDeviceToHost, stream),
                      "mrc: copy to the host");
            check_hip(hipEventRecord(sg.done[j], stream), "mrc: record");
            pending_bytes[j] = bytes;
            flush(j ^ 1);   // write the previous chunk while this one is encoded
        }
        flush(0);
        flush(1);
        if (std::fflush(file.f) != 0) fail(NOA_HIP_ERROR_RUNTIME, "Failed to write the file. %s", std::strerror(errno));
        return NOA_HIP_OK;
    } catch (const Fail& f) {
        err = f.msg;
        return f.code;
    }
}

} // namespace noa_hip
