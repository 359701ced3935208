// Test program of the C++ facade (include/noa_hip.hpp).
//
//   test_facade          descriptors built from noa-style arguments, and validation errors (no GPU needed)
//   test_facade --gpu    identity-pose extraction and insertion through the facade on device 0 (closed forms)
//
// Prints one line per check and exits non-zero on the first failure.
#include "noa_hip.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace nx = noa_hip::xform;
using noa_hip::c32;
using noa_hip::Empty;
using noa_hip::Mat33;
using noa_hip::Quaternion;
using noa_hip::Shape4;
using noa_hip::View;

#define EXPECT(cond)                                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond);              \
            std::exit(1);                                                            \
        }                                                                            \
    } while (0)

#define HIP_OK(x) EXPECT((x) == hipSuccess)

static void check_layouts() {
    EXPECT(nx::Layout("h2h").bits == NOA_HIP_H2H);
    EXPECT(nx::Layout("h").bits == NOA_HIP_H2H);
    EXPECT(nx::Layout("hc2h").bits == NOA_HIP_HC2H);
    EXPECT(nx::Layout("h2hc").bits == NOA_HIP_H2HC);
    EXPECT(nx::Layout("hc2hc").bits == NOA_HIP_HC2HC);
    EXPECT(nx::Layout("hc").bits == NOA_HIP_HC2HC);
    EXPECT(nx::Layout("f2f").bits == NOA_HIP_F2F && nx::Layout("fc2fc").bits == NOA_HIP_FC2FC);
    EXPECT(nx::Layout("fc2f").bits == NOA_HIP_FC2F && nx::Layout("f2fc").bits == NOA_HIP_F2FC);
    std::printf("ok layouts\n");
}

static void check_descriptor() {
    // An insertion of 5 slices 64x64 (rfft 64x33, padded rows of 40) into a 64^3 volume, per-slice inverse
    // rotations on the device (here a fake device address: nothing is dereferenced), constant slice weight.
    auto* slices = reinterpret_cast<c32*>(0x1000);
    auto* volume = reinterpret_cast<c32*>(0x2000);
    auto* rot = reinterpret_cast<Mat33<float>*>(0x3000);
    const View<c32> s(slices, Shape4(5, 1, 64, 33), Shape4(64 * 40, 64 * 40, 40, 1));
    const View<c32> v(volume, Shape4(1, 64, 64, 33));
    const View<Mat33<float>> r(rot, Shape4(1, 1, 1, 5));
    nx::InsertCentralSlicesOptions opt;
    opt.interp = nx::Interp::NEAREST;
    opt.windowed_sinc = {0.02, 0.05};
    opt.fftfreq_cutoff = 0.4;
    opt.target_shape = Shape4(1, 128, 128, 128);
    const auto a = nx::details::make_args(nx::Layout("hc2h").bits, s, 2.5f, Shape4(5, 1, 64, 64), v, Empty{},
                                          Shape4(1, 64, 64, 64), Empty{}, r, opt);
    EXPECT(a.remap == NOA_HIP_HC2H);
    EXPECT(a.interp == NOA_HIP_INTERP_NEAREST);
    EXPECT(a.input.data == slices && a.input.dtype == NOA_HIP_C32);
    EXPECT(a.input.shape[0] == 5 && a.input.shape[3] == 33 && a.input.strides[2] == 40 && a.input.strides[0] == 2560);
    EXPECT(a.input_weight.is_value == 1 && a.input_weight.dtype == NOA_HIP_F32 && a.input_weight.value[0] == 2.5);
    EXPECT(a.output.data == volume && a.output.strides[1] == 64 * 33 && a.output.strides[3] == 1);
    EXPECT(a.output_weight.data == nullptr && a.output_weight.is_value == 0);
    EXPECT(a.input_shape[3] == 64 && a.output_shape[1] == 64);
    EXPECT(a.target_shape[1] == 128);
    EXPECT(a.rotation.kind == NOA_HIP_MAT33_F32 && a.rotation.batched == 1 && a.rotation.count == 5 &&
           a.rotation.data == rot);
    EXPECT(a.scaling.kind == NOA_HIP_XFORM_NONE);
    EXPECT(a.fftfreq_cutoff == 0.4 && a.fftfreq_sinc == 0.02 && a.fftfreq_blackman == 0.05);

    // One broadcast quaternion: a host value.
    const Quaternion<double> q{0, 0, 0, 1};
    const auto b = nx::details::make_args(NOA_HIP_HC2HC, v, Empty{}, Shape4(1, 64, 64, 64), s, Empty{},
                                          Shape4(5, 1, 64, 64), Empty{}, q, nx::ExtractCentralSlicesOptions{});
    EXPECT(b.rotation.kind == NOA_HIP_QUAT_F64 && b.rotation.batched == 0 && b.rotation.data == &q);
    EXPECT(b.interp == NOA_HIP_INTERP_LINEAR && b.fftfreq_cutoff == 0.5 && b.fftfreq_sinc == -1);
    std::printf("ok descriptors\n");
}

static void check_errors() {
    // Wrong volume shape for the array: rejected before anything touches the (fake) memory.
    const View<c32> s(reinterpret_cast<c32*>(0x10000000), Shape4(2, 1, 32, 17));
    const View<c32> v(reinterpret_cast<c32*>(0x20000000), Shape4(1, 32, 32, 17));
    const Mat33<float> eye{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    bool thrown = false;
    try {
        nx::extract_central_slices_3d<"hc2hc">(v, Empty{}, Shape4(1, 16, 32, 32), s, Empty{}, Shape4(2, 1, 32, 32),
                                             Empty{}, eye);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "shape") != nullptr);
    }
    EXPECT(thrown);
    // An interpolation method outside noa's Interp values.
    thrown = false;
    try {
        nx::ExtractCentralSlicesOptions opt;
        opt.interp = static_cast<nx::Interp>(9);
        nx::extract_central_slices_3d<"hc2hc">(v, Empty{}, Shape4(1, 32, 32, 32), s, Empty{}, Shape4(2, 1, 32, 32),
                                             Empty{}, eye, opt);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "nterp") != nullptr);
    }
    EXPECT(thrown);
}

// Identity poses on device 0: the extracted slice is the central plane z = 0 of the volume (exactly), and the
// inserted slice lands on the plane z = 0 multiplied by the window at 0 (noa's windowed_sinc(0)).
static void check_gpu() {
    const int n = 32, wr = n / 2 + 1;
    const size_t nvox = static_cast<size_t>(n) * n * wr, npix = static_cast<size_t>(n) * wr;
    std::vector<c32> hv(nvox);
    for (size_t i = 0; i < nvox; ++i) hv[i] = c32(static_cast<float>(i % 97) - 40.f, static_cast<float>(i % 31));
    c32 *dv = nullptr, *ds = nullptr, *dins = nullptr;
    float* dw = nullptr;
    HIP_OK(hipMalloc(&dv, nvox * sizeof(c32)));
    HIP_OK(hipMalloc(&ds, npix * sizeof(c32)));
    HIP_OK(hipMalloc(&dins, nvox * sizeof(c32)));
    HIP_OK(hipMalloc(&dw, nvox * sizeof(float)));
    HIP_OK(hipMemcpy(dv, hv.data(), nvox * sizeof(c32), hipMemcpyHostToDevice));
    HIP_OK(hipMemset(dins, 0, nvox * sizeof(c32)));
    HIP_OK(hipMemset(dw, 0, nvox * sizeof(float)));

    const Mat33<float> eye{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    const View<c32> vol(dv, Shape4(1, n, n, wr));
    const View<c32> sl(ds, Shape4(1, 1, n, wr));
    nx::extract_central_slices_3d<"h2h">(vol, Empty{}, Shape4(1, n, n, n), sl, Empty{}, Shape4(1, 1, n, n), Empty{},
                                         eye);
    HIP_OK(hipDeviceSynchronize());
    std::vector<c32> hs(npix);
    HIP_OK(hipMemcpy(hs.data(), ds, npix * sizeof(c32), hipMemcpyDeviceToHost));
    // Non-centred volume: plane z = 0 is index 0. Pixels beyond the 0.5 cutoff are 0.
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < wr; ++x) {
            const int fy = y < (n + 1) / 2 ? y : y - n;
            const double f2 = (static_cast<double>(fy) / n) * (static_cast<double>(fy) / n) +
                              (static_cast<double>(x) / n) * (static_cast<double>(x) / n);
            const c32 want = f2 > 0.25 ? c32(0, 0) : hv[static_cast<size_t>(y) * wr + x];
            EXPECT(hs[static_cast<size_t>(y) * wr + x] == want);
        }
    std::printf("ok gpu extraction (identity pose, exact)\n");

    const View<c32> ins(dins, Shape4(1, n, n, wr));
    const View<float> w(dw, Shape4(1, n, n, wr));
    nx::insert_central_slices_3d<"h2h">(sl, Empty{}, Shape4(1, 1, n, n), ins, w, Shape4(1, n, n, n), Empty{}, eye);
    HIP_OK(hipDeviceSynchronize());
    std::vector<c32> hi(nvox);
    std::vector<float> hw(nvox);
    HIP_OK(hipMemcpy(hi.data(), dins, nvox * sizeof(c32), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(hw.data(), dw, nvox * sizeof(float), hipMemcpyDeviceToHost));
    const float w0 = 0.99999994f;   // noa's windowed_sinc(0) in FP32: sinc(0) * (0.42 + 0.5 + 0.08)
    double max_err = 0;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < wr; ++x) {
                const size_t i = (static_cast<size_t>(z) * n + y) * wr + x;
                const c32 want = z == 0 ? hs[static_cast<size_t>(y) * wr + x] * w0 : c32(0, 0);
                max_err = std::fmax(max_err, std::abs(hi[i] - want));
                if (z != 0) EXPECT(std::fabs(hw[i]) < 1e-6f);
            }
    std::printf("gpu insertion max abs error %.3g\n", max_err);
    EXPECT(max_err < 1e-4);
    std::printf("ok gpu insertion (identity pose, plane z = 0)\n");
    // r2c of a 16x16 image (norm FORWARD: the DC term is the mean), then c2r back to the image.
    {
        const int m = 16, mr = m / 2 + 1;
        float* dr = nullptr;
        c32* dc = nullptr;
        HIP_OK(hipMalloc(&dr, m * m * sizeof(float)));
        HIP_OK(hipMalloc(&dc, m * mr * sizeof(c32)));
        std::vector<float> hr(m * m);
        double mean = 0;
        for (int i = 0; i < m * m; ++i) { hr[i] = static_cast<float>((i * 37) % 11) - 5.f; mean += hr[i]; }
        mean /= m * m;
        HIP_OK(hipMemcpy(dr, hr.data(), hr.size() * sizeof(float), hipMemcpyHostToDevice));
        const View<float> rv(dr, Shape4(1, 1, m, m));
        const View<c32> cv(dc, Shape4(1, 1, m, mr));
        noa_hip::fft::r2c(rv, cv);
        HIP_OK(hipDeviceSynchronize());
        c32 dc0;
        HIP_OK(hipMemcpy(&dc0, dc, sizeof(c32), hipMemcpyDeviceToHost));
        EXPECT(std::abs(dc0.real() - mean) < 1e-5 && std::abs(dc0.imag()) < 1e-5);
        noa_hip::fft::c2r(cv, rv);
        HIP_OK(hipDeviceSynchronize());
        std::vector<float> back(m * m);
        HIP_OK(hipMemcpy(back.data(), dr, back.size() * sizeof(float), hipMemcpyDeviceToHost));
        double ferr = 0;
        for (int i = 0; i < m * m; ++i) ferr = std::fmax(ferr, std::fabs(back[i] - hr[i]));
        EXPECT(ferr < 1e-5);
        std::printf("ok gpu fft round trip (max error %.3g)\n", ferr);
        HIP_OK(hipFree(dr));
        HIP_OK(hipFree(dc));
    }
    // An MRC round trip through the device: f32 written as I16, read back into f32.
    {
        const int n = 8 * 6 * 10;
        float* dm = nullptr;
        HIP_OK(hipMalloc(&dm, n * sizeof(float)));
        std::vector<float> hm(n);
        for (int i = 0; i < n; ++i) hm[i] = static_cast<float>(i % 200) - 100.f;
        HIP_OK(hipMemcpy(dm, hm.data(), n * sizeof(float), hipMemcpyHostToDevice));
        const View<float> mv(dm, Shape4(1, 8, 6, 10));
        noa_hip::io::WriteOption wo;
        wo.spacing = {1.5, 1.5, 1.5};
        wo.dtype = noa_hip::io::DataType::I16;
        noa_hip::io::write_image(mv, "test_facade_io.mrc", wo);
        HIP_OK(hipMemset(dm, 0, n * sizeof(float)));
        const noa_hip_mrc_header h = noa_hip::io::read_header("test_facade_io.mrc");
        EXPECT(h.shape[1] == 8 && h.shape[2] == 6 && h.shape[3] == 10 && h.dtype == NOA_HIP_IO_I16);
        noa_hip::io::read_image("test_facade_io.mrc", mv);
        std::vector<float> back(n);
        HIP_OK(hipMemcpy(back.data(), dm, n * sizeof(float), hipMemcpyDeviceToHost));
        EXPECT(back == hm);
        std::remove("test_facade_io.mrc");
        std::printf("ok gpu mrc round trip\n");
        HIP_OK(hipFree(dm));
    }
    // Prefilter of a constant (1, 8, 16, 32) volume in place: constant coefficients, up to the truncated causal
    // start (|z|^9 of the value along the 8-sample axis).
    {
        const size_t np = 8 * 16 * 32;
        float* dp = nullptr;
        HIP_OK(hipMalloc(&dp, np * sizeof(float)));
        std::vector<float> hp(np, 3.f);
        HIP_OK(hipMemcpy(dp, hp.data(), np * sizeof(float), hipMemcpyHostToDevice));
        const View<float> pv(dp, Shape4(1, 8, 16, 32));
        nx::cubic_bspline_prefilter(pv, pv);
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipMemcpy(hp.data(), dp, np * sizeof(float), hipMemcpyDeviceToHost));
        double perr = 0;
        for (float v : hp) perr = std::fmax(perr, std::fabs(v - 3.0));
        EXPECT(perr < 1e-4);
        std::printf("ok gpu prefilter (max deviation %.3g)\n", perr);
        HIP_OK(hipFree(dp));
    }
    HIP_OK(hipFree(dv));
    HIP_OK(hipFree(ds));
    HIP_OK(hipFree(dins));
    HIP_OK(hipFree(dw));
}

static void check_phase_shift_errors() {
    // A depth-2 shape is not a 2d phase shift: rejected before any memory is touched.
    const View<c32> s(reinterpret_cast<c32*>(0x10000000), Shape4(2, 2, 32, 17));
    bool thrown = false;
    try {
        noa_hip::signal::phase_shift_2d<"h2h">(Empty{}, s, Shape4(2, 2, 32, 32), nx::Vec2<float>{1.f, 2.f});
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
    }
    EXPECT(thrown);
}

static void check_insert_extract_descriptor() {
    // 20 input slices (inverse rotations on the device) -> 5 output slices (one broadcast forward rotation).
    auto* in = reinterpret_cast<c32*>(0x10000000);
    auto* out = reinterpret_cast<c32*>(0x20000000);
    auto* wout = reinterpret_cast<float*>(0x30000000);
    auto* rot = reinterpret_cast<Mat33<float>*>(0x40000000);
    const View<c32> i(in, Shape4(20, 1, 64, 33));
    const View<c32> o(out, Shape4(5, 1, 64, 33));
    const View<float> w(wout, Shape4(5, 1, 64, 33));
    const View<Mat33<float>> r(rot, Shape4(1, 1, 1, 20));
    const Mat33<float> eye{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    nx::InsertAndExtractCentralSlicesOptions opt;
    opt.input_windowed_sinc = {0.01, 0.02};
    opt.w_windowed_sinc = {0.03, 0.04};
    opt.add_to_output = true;
    opt.fftfreq_cutoff = 0.8;
    const auto a = nx::details::make_insert_extract_args(nx::Layout("hc2hc").bits, i, Empty{}, Shape4(20, 1, 64, 64),
                                                         o, w, Shape4(5, 1, 64, 64), Empty{}, r, Empty{}, eye, opt);
    EXPECT(a.remap == NOA_HIP_HC2HC && a.interp == NOA_HIP_INTERP_LINEAR);
    EXPECT(a.input.data == in && a.output.data == out && a.output_weight.data == wout);
    EXPECT(a.input_weight.data == nullptr && a.input_weight.is_value == 0);
    EXPECT(a.input_rotation.batched == 1 && a.input_rotation.count == 20 && a.input_rotation.data == rot);
    EXPECT(a.output_rotation.batched == 0 && a.output_rotation.data == &eye);
    EXPECT(a.input_scaling.kind == NOA_HIP_XFORM_NONE && a.output_scaling.kind == NOA_HIP_XFORM_NONE);
    EXPECT(a.input_fftfreq_sinc == 0.01 && a.input_fftfreq_blackman == 0.02);
    EXPECT(a.w_fftfreq_sinc == 0.03 && a.w_fftfreq_blackman == 0.04);
    EXPECT(a.add_to_output == 1 && a.correct_weights == 0 && a.fftfreq_cutoff == 0.8);
    // correct_weights with add_to_output is refused before any memory is touched (FourierProject.hpp:2011-2018).
    opt.correct_weights = true;
    bool thrown = false;
    try {
        nx::insert_and_extract_central_slices_3d<"hc2hc">(i, Empty{}, Shape4(20, 1, 64, 64), o, w,
                                                          Shape4(5, 1, 64, 64), Empty{}, r, Empty{}, eye, opt);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "correct_weights") != nullptr);
    }
    EXPECT(thrown);
    std::printf("ok insert_extract descriptor\n");
}

static void check_bspline_prefilter_errors() {
    // An input that does not broadcast into the output is refused before any memory is touched.
    const View<float> in(reinterpret_cast<float*>(0x10000000), Shape4(1, 4, 3, 8));
    const View<float> out(reinterpret_cast<float*>(0x20000000), Shape4(1, 4, 4, 8));
    bool thrown = false;
    try {
        nx::cubic_bspline_prefilter(in, out);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "Cannot broadcast") != nullptr);
    }
    EXPECT(thrown);
}

static void check_fft_errors() {
    // noa's r2c shape rule (fft/Transform.hpp:169-173) is checked before any memory is touched.
    const View<float> real(reinterpret_cast<float*>(0x10000000), Shape4(1, 1, 16, 16));
    const View<c32> bad(reinterpret_cast<c32*>(0x20000000), Shape4(1, 1, 16, 8));
    bool thrown = false;
    try {
        noa_hip::fft::r2c(real, bad);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "non-redundant shape") != nullptr);
    }
    EXPECT(thrown);
}

static void check_ctf_object() {
    // noa's constructor: 300 kV -> lambda = 0.0196875 A, k1 = pi * lambda, k3 = atan(A / sqrt(1 - A^2)), defocus in
    // angstroms with the sign flipped, cs in angstroms, bfactor / 4.
    const noa_hip::signal::CTFAnisotropic<double> c({1.0, 1.1}, {1.5, 0.05, 0.3}, 300.0, 0.07, 2.7, 0.2, -20.0, 2.0);
    EXPECT(std::fabs(c.k1 / 3.141592653589793 - 0.0196875) < 1e-7);
    EXPECT(std::fabs(c.k3 - std::asin(0.07)) < 1e-15);
    EXPECT(c.defocus_value == -15000.0 && c.defocus_astigmatism == -500.0 && c.defocus_angle == 0.3);
    EXPECT(c.cs == 2.7e7 && c.voltage == 3e5 && c.bfactor_forth == -5.0 && c.scale == 2.0);
    EXPECT(c.pixel_size[1] == 1.1 && c.phase_shift == 0.2);
    const double lam = c.k1 / 3.141592653589793;
    EXPECT(std::fabs(c.k2 - 3.141592653589793 * 0.5 * 2.7e7 * lam * lam * lam) < 1e-9 * c.k2);
    const noa_hip::signal::CTFAnisotropic<float> f({1.f, 1.f}, {1.f, 0.f, 0.f}, 300.f, 0.1f, 2.7f);
    static_assert(sizeof(f) == 56 && alignof(noa_hip::signal::CTFAnisotropic<double>) == 16);
    const auto d = noa_hip::signal::details::ctf(f);
    EXPECT(d.dtype == NOA_HIP_F32 && d.batched == 0 && d.data == &f);
    const noa_hip::signal::CTFIsotropic<double> i(1.2, 1.5, 300.0, 0.07, 2.7, 0.2, -20.0, 2.0);
    EXPECT(i.defocus == -15000.0 && i.cs == 2.7e7 && i.voltage == 3e5 && i.bfactor_forth == -5.0);
    EXPECT(i.k1 == c.k1 && i.k2 == c.k2 && i.k3 == c.k3 && i.pixel_size == 1.2 && i.scale == 2.0);
    static_assert(sizeof(noa_hip::signal::CTFIsotropic<float>) == 44);
    const auto di = noa_hip::signal::details::ctf(i);
    EXPECT(di.dtype == NOA_HIP_F64 && di.batched == 0);
    // ctf_check_parameters (signal/CTF.hpp:143-189): the output's shape is checked before any memory is touched.
    const View<float> out(reinterpret_cast<float*>(0x10000000), Shape4(2, 8, 8, 8));
    bool thrown = false;
    try {
        noa_hip::signal::ctf_isotropic<"h2h">(View<float>{}, out, Shape4(2, 8, 8, 8), i);
    } catch (const noa_hip::Exception& e) {
        thrown = true;
        std::printf("ok error: %s\n", e.what());
        EXPECT(std::strstr(e.what(), "expected shape") != nullptr);
    }
    EXPECT(thrown);
    std::printf("ok ctf object\n");
}

int main(int argc, char** argv) {
    check_layouts();
    check_descriptor();
    check_errors();
    check_phase_shift_errors();
    check_insert_extract_descriptor();
    check_ctf_object();
    check_bspline_prefilter_errors();
    check_fft_errors();
    if (argc > 1 && std::strcmp(argv[1], "--gpu") == 0) check_gpu();
    std::printf("all facade checks passed\n");
    return 0;
}
