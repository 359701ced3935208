#!/usr/bin/env python3
"""Throughput of the central-slice hot path on one MI355X: 256^3 extraction + insertion, in slices/s.

One step = extract 4096 slices (256x129 complex64, "hc") from a 256^3 volume, then insert 10,000 CTF-weighted
slices (with CTF^2 weights) into a 256^3 volume + weights (FP32, LINEAR, windowed sinc of 1 voxel). Inputs
are synthetic (seeded) and already resident in HBM when the timed region starts.

Prints ONE JSON line (stdout): the metric, per-kernel achieved bandwidth against the MI355X HBM peak, and the
CPU baseline (the restated noa CPU path, timed on a bounded sample on this host).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

# HBM-side bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over
# `bench.py --probe`, scaled by the same counters on a 1 GiB device copy (fetch x 2.0, write x 1.0); recorded in
# profiles/insert_wave_kernel_256cube_10k_slices.md and profiles/extract_kernel_256cube_4096_slices.md for the
# default workload sizes.
MEASURED_TRAFFIC = {
    ("insert", 10_000): (8_572_515 * 2.0 + 387_096) * 1024,
    # the gather kernel plus the block copy of the volume it reads (counters measured with 4x2x2 blocks)
    ("extract", 4096): (1_868_633 * 2.0 + 1_056_768 + 33_053 * 2.0 + 68_354) * 1024,
}
METRIC = "slices/s and % HBM peak, 256\u00b3 insert+extract; p50/p99 request latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes_extract(b, n):
    """Output slices + volume + one Mat33 payload and one (y, x) f32 shift per slice (each read or written once);
    the phase shift is fused, so the slices are written once."""
    return b * n * (n // 2 + 1) * 8 + n * n * (n // 2 + 1) * 8 + b * 36 + b * 8


def algorithmic_bytes_insert(b, n):
    """c32 slices + f32 weights read once, c32 volume and f32 weights read and written once, Mat33 per slice."""
    px = n * (n // 2 + 1)
    vox = n * n * (n // 2 + 1)
    return b * px * (8 + 4) + 2 * vox * 8 + 2 * vox * 4 + b * 36


def cpu_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))   # at most the CPU share of one GPU box


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(ext, ins, n_ext_sample=64, n_ins_sample=48):
    """The oracle (noa's CPU algorithm, OpenMP) on a bounded sample of the same workload.

    Threads follow noa's heuristic min(cores, n_elements / 2^20) evaluated on the FULL workload
    (cpu/Iwise.hpp:104-116): all cores for the 4096-slice extraction, 8 for the 256^3 insertion."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle as O

    cores = cpu_cores()
    n = ext.volume.shape[1]
    vol = ext.volume.cpu().numpy()
    rot = ext.poses[:n_ext_sample]
    out = np.zeros((n_ext_sample, 1, n, n // 2 + 1), np.complex64)
    t_ext = min(cores, (4096 * n * (n // 2 + 1)) >> 20)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), out, None, (n_ext_sample, 1, n, n), None, rot,
                                    n_threads=t_ext, force_threads=True)
        O.phase_shift_2d("hc2hc", out, out, (n_ext_sample, 1, n, n), ext.shifts[:n_ext_sample])
        times.append(time.perf_counter() - t0)
    ext_per_slice = statistics.median(times) / n_ext_sample

    s = ins.slices[:n_ins_sample].cpu().numpy()
    w = ins.weights[:n_ins_sample].cpu().numpy()
    inv = np.ascontiguousarray(np.transpose(ins.poses[:n_ins_sample], (0, 2, 1)))
    v = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    wv = np.zeros((1, n, n, n // 2 + 1), np.float32)
    t_ins = min(cores, (n * n * (n // 2 + 1)) >> 20)
    t0 = time.perf_counter()
    O.insert_central_slices_3d("h2h", s, w, (n_ins_sample, 1, n, n), v, wv, (1, n, n, n), None, inv,
                               n_threads=t_ins, force_threads=True)
    ins_per_slice = (time.perf_counter() - t0) / n_ins_sample

    n_e, n_i = ext.slice_shape[0], ins.slice_shape[0]
    step_s = ext_per_slice * n_e + ins_per_slice * n_i
    return {
        "value": (n_e + n_i) / step_s,
        "unit": "slices/s",
        "cores": max(t_ext, t_ins),
        "kind": "port",
        "sample": (f"oracle (noa CPU algorithm restated, OpenMP): extraction + phase shift of {n_ext_sample} of the {n_e} poses "
                   f"with {t_ext} threads ({ext_per_slice * 1e3:.3f} ms/slice), insertion of {n_ins_sample} of the "
                   f"{n_i} slices into 256^3 with {t_ins} threads ({ins_per_slice * 1e3:.1f} ms/slice), "
                   f"extrapolated linearly to one step; CPU: {cpu_model()}"),
    }


def serving_report(torch, dev, n_requests, qps, budget_gb=16.0):
    """Poisson stream of mixed-box extraction/insertion requests through the continuous-batching loop
    (noa_amd.serve): p50/p99 latency and served slices/s. Request inputs are on the device before the stream
    starts; latency is completion (event passed) minus arrival."""
    from noa_amd import serve, workloads

    vols = workloads.serving_volumes(dev)
    reqs = workloads.serving_requests(dev, n_requests=n_requests, qps=qps)
    server = serve.Server(vols, budget_bytes=int(budget_gb * (1 << 30)), n_streams=3)
    # Warm-up: the insertion work plans of every box and power-of-two batch (built once per geometry).
    for box in workloads.SERVING_BOXES:
        for k in (64, 128, 256, 512):
            r = serve.Request(rid=-1, kind="insert", box=box, rotations=torch.eye(3, device=dev).expand(k, 3, 3).contiguous(),
                              slices=torch.zeros((k, 1, box, box // 2 + 1), dtype=torch.complex64, device=dev),
                              weights=torch.zeros((k, 1, box, box // 2 + 1), dtype=torch.float32, device=dev))
            serve.run_alone(server, r)
    server.stats = serve.ServeStats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = serve.run_open_loop(server, reqs, timeout_s=300)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kinds = [r.kind for r in reqs]
    return {
        "p50_ms": stats.percentile(50),
        "p99_ms": stats.percentile(99),
        "qps": qps,
        "requests": len(reqs),
        "extract_requests": kinds.count("extract"),
        "insert_requests": kinds.count("insert"),
        "slices_per_s": stats.slices / wall,
        "launches": stats.launches,
        "merged_requests": stats.merged,
        "wall_s": wall,
        "mix": "box {64,128,192,256,384} p {.3,.3,.2,.15,.05}; extract K~U{1..512} from 4 resident volumes per box "
               "or insert K CTF-weighted slices into a per-request accumulator; Poisson arrivals",
    }


def precision_report(ext, nx, torch, n_ref=128):
    """Extraction from the 256^3 volume stored as FP32, FP16 pairs and BF16 pairs (FP32 math): rel-L2 and max error
    of each against the FP64 oracle (double volume and rotations) on the first n_ref poses, and the time of each for
    all poses."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import oracle as O

    n = ext.volume.shape[1]
    b = ext.slice_shape[0]
    vol64 = ext.volume.cpu().numpy().astype(np.complex128)
    rot64 = ext.poses[:n_ref].astype(np.float64)
    ref = np.zeros((n_ref, 1, n, n // 2 + 1), np.complex128)
    O.extract_central_slices_3d("hc2hc", vol64, None, (1, n, n, n), ref, None, (n_ref, 1, n, n), None, rot64,
                                n_threads=cpu_cores(), force_threads=True)
    storages = {
        "fp32": ext.volume,
        "fp16": ext.volume.to(torch.complex32),
        "bf16": nx.ComplexBF16.from_complex(ext.volume),
    }
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=ext.volume.device)
    report = {"report": "extraction volume-storage precision", "config": "256^3 hc volume CN(0,1)*exp(-20|f|^2), "
              f"{b} uniform poses (error on the first {n_ref}), LINEAR, cutoff 0.5, FP32 math",
              "oracle": "FP64 restatement of noa's CPU extraction (double volume and rotations)"}
    for name, v in storages.items():
        def run():
            nx.extract_central_slices_3d("hc2hc", v, None, ext.volume_shape, out, None, ext.slice_shape, None,
                                         ext.rotations)
        run()
        torch.cuda.synchronize()
        got = out[:n_ref].cpu().numpy().astype(np.complex128)
        err = got - ref
        # Pixels on the cutoff circle (within 1e-3) can fall on either side in FP32 vs FP64 coordinates.
        fy = (np.arange(n) - n // 2)[:, None] / n
        fx = np.arange(n // 2 + 1)[None, :] / n
        inner = np.broadcast_to(np.abs(np.sqrt(fy ** 2 + fx ** 2) - 0.5) > 1e-3, err.shape)
        for _ in range(3):   # warm: the first calls of a storage also ramp the clocks and grow the memory pool
            run()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        t0.record()
        for _ in range(20):
            run()
        t1.record()
        torch.cuda.synchronize()
        report[name] = {
            "rel_l2": float(np.linalg.norm(err) / np.linalg.norm(ref)),
            "rel_l2_off_cutoff_circle": float(np.linalg.norm(err[inner]) / np.linalg.norm(ref[inner])),
            "max_abs": float(np.abs(err).max()),
            "max_abs_over_max_ref": float(np.abs(err).max() / np.abs(ref).max()),
            "volume_bytes": int(v.pairs.numel() * 2 if name == "bf16" else v.numel() * v.element_size()),
            "ms": t0.elapsed_time(t1) / 20,
        }
        log(f"{name}: rel-L2 {report[name]['rel_l2']:.3g}, max abs {report[name]['max_abs']:.3g}, "
            f"{report[name]['ms']:.3f} ms")
    return report


def neighbours_report(torch, dev, nx, ins, n, reps=5):
    """The steps on either side of the hot path, timed alone (not part of the step): the CTF prologue of the
    insertion (slice * CTF and CTF^2 in one pass over the insertion workload's raw spectra) and an extraction from a
    virtual volume (a 61-view +-60 degree tilt series of 256^2 slices reprojected at 61 tilts offset by 1 degree)."""
    import numpy as np
    import noa_amd.signal as ns
    rep = {}
    b = ins.raw.shape[0]
    vals = torch.empty_like(ins.raw)
    wts = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.float32, device=dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) * 1e-3 / reps

    t = timed(lambda: ns.ctf_anisotropic_weighted("h2h", ins.raw, vals, wts, ins.slice_shape, ins.ctfs))
    nbytes = b * n * (n // 2 + 1) * (8 + 8 + 4) + b * 56
    dv = (vals - ins.slices).abs().max().item() / ins.slices.abs().max().item()
    rep["ctf_prologue"] = {
        "workload": f"{b} slices 256^2 (h2h): value * CTF and CTF^2, CTFAnisotropic<f32> per slice",
        "avg_ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
        "frac_hbm_peak": nbytes / t / 1e9 / HBM_PEAK_GBS, "slices_per_s": b / t,
        "max_rel_diff_vs_workload_inputs": dv}
    log(f"ctf prologue: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s")
    del vals, wts

    tilts_in = np.deg2rad(np.arange(-60, 61, 2, dtype=np.float64))
    tilts_out = tilts_in + np.deg2rad(1.0)
    rin = np.stack([nx.euler2matrix([0.0, -a, 0.0], axes="zyx", dtype=np.float32) for a in tilts_in])
    rout = np.stack([nx.euler2matrix([0.0, a, 0.0], axes="zyx", dtype=np.float32) for a in tilts_out])
    b_in, b_out = len(tilts_in), len(tilts_out)
    src = ins.raw[:b_in]
    out = torch.empty((b_out, 1, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    wout = torch.empty((b_out, 1, n, n // 2 + 1), dtype=torch.float32, device=dev)
    rin_d, rout_d = torch.from_numpy(rin).to(dev), torch.from_numpy(rout).to(dev)
    t = timed(lambda: nx.insert_and_extract_central_slices_3d(
        "h2h", src, None, (b_in, 1, n, n), out, wout, (b_out, 1, n, n), None, rin_d, None, rout_d))
    pairs = b_out * n * (n // 2 + 1) * b_in
    rep["virtual_volume"] = {
        "workload": f"insert_and_extract_central_slices_3d: {b_in} tilts (+-60, 2 deg) of 256^2 -> {b_out} tilts "
                    "offset by 1 deg, h2h, LINEAR, default windows",
        "avg_ms": t * 1e3, "output_slices_per_s": b_out / t, "pixel_slice_pairs_per_s": pairs / t,
        "nonzero_weight_fraction": (wout != 0).float().mean().item()}
    log(f"virtual volume: {t * 1e3:.3f} ms, {pairs / t / 1e9:.1f} G pixel-slice pairs/s")
    del out, wout

    # Isotropic CTF of a reconstructed 256^3 rfft volume (h2h): the volume multiplied by CTF(|f|) in place.
    spec = torch.randn((1, n, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    ciso = ns.CTFIsotropic(1.0, 1.5, 300, 0.07, 2.7)
    t = timed(lambda: ns.ctf_isotropic("h2h", spec, spec, (1, n, n, n), ciso))
    nbytes = 2 * spec.numel() * 8
    rep["ctf_isotropic_volume"] = {
        "workload": f"ctf_isotropic of a {n}^3 c32 rfft volume (h2h) in place, CTFIsotropic<f32>",
        "avg_ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
        "frac_hbm_peak": nbytes / t / 1e9 / HBM_PEAK_GBS}
    log(f"ctf isotropic {n}^3: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s")
    del spec

    # Cubic B-spline prefilter of a 256^3 f32 volume in place (before CUBIC_BSPLINE interpolation): three passes
    # (x, y, z), each reading and writing the volume once.
    vol = torch.randn((1, n, n, n), device=dev, dtype=torch.float32)
    t = timed(lambda: nx.cubic_bspline_prefilter(vol, vol))
    nbytes = 3 * 2 * vol.numel() * 4
    rep["bspline_prefilter"] = {
        "workload": f"cubic_bspline_prefilter of a {n}^3 f32 volume in place (x, y, z passes)",
        "avg_ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
        "frac_hbm_peak": nbytes / t / 1e9 / HBM_PEAK_GBS}
    log(f"bspline prefilter: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s")
    del vol

    # FFTs on either side of the path (rocFFT plans cached as in noa): the c2r of the reconstructed 256^3 rfft
    # volume, and the r2c of a stack of 1024 real 256^2 images into "h" slices.
    import noa_amd.fft as nf
    spec = torch.randn((1, n, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    real = torch.empty((1, n, n, n), dtype=torch.float32, device=dev)
    t = timed(lambda: nf.c2r(spec, real))
    nbytes = spec.numel() * 8 + real.numel() * 4
    rep["fft_c2r_volume"] = {
        "workload": f"fft::c2r of a {n}^3 volume (c32 rfft -> f32), norm FORWARD",
        "avg_ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
        "frac_hbm_peak": nbytes / t / 1e9 / HBM_PEAK_GBS}
    log(f"fft c2r {n}^3: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s")
    del spec, real
    imgs = torch.randn((1024, 1, n, n), dtype=torch.float32, device=dev)
    out = torch.empty((1024, 1, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    t = timed(lambda: nf.r2c(imgs, out))
    nbytes = imgs.numel() * 4 + out.numel() * 8
    rep["fft_r2c_slices"] = {
        "workload": f"fft::r2c of 1024 real {n}^2 images (f32 -> c32 rfft), norm FORWARD (1/N fused in the transform)",
        "avg_ms": t * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": nbytes / t / 1e9,
        "frac_hbm_peak": nbytes / t / 1e9 / HBM_PEAK_GBS, "slices_per_s": 1024 / t}
    log(f"fft r2c 1024 x {n}^2: {t * 1e3:.3f} ms, {nbytes / t / 1e9:.0f} GB/s")

    # MRC IO: a stack of 1024 256^2 images stored as I16 (the usual raw-data mode), written from and read back
    # into f32 device memory (chunked through pinned buffers, decoded and encoded on the GPU). The file sits in
    # the page cache after the write, so the read rate is host memory + PCIe, not a disk's.
    import tempfile
    import noa_amd.io as nio
    stack = torch.round(imgs * 100).contiguous()
    del out
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "stack.mrc")
        t0 = time.perf_counter()
        nio.write_image(stack, path, spacing=(1.0, 1.0, 1.0), dtype="I16")
        t_w = time.perf_counter() - t0
        dst = torch.empty_like(stack)
        nio.read_image(path, output=dst)
        t0 = time.perf_counter()
        nio.read_image(path, output=dst)
        t_r = time.perf_counter() - t0
        ok = bool(torch.equal(dst, torch.clamp(stack, -32768, 32767)))
    fbytes = stack.numel() * 2
    rep["mrc_io"] = {
        "workload": f"1024 x {n}^2 f32 device images <-> MRC mode 1 (I16) file, page-cached",
        "read_ms": t_r * 1e3, "write_ms": t_w * 1e3, "file_bytes": fbytes,
        "read_file_GBs": fbytes / t_r / 1e9, "write_file_GBs": fbytes / t_w / 1e9, "round_trip_exact": ok}
    log(f"mrc read {t_r * 1e3:.1f} ms ({fbytes / t_r / 1e9:.1f} GB/s of file), write {t_w * 1e3:.1f} ms")
    del imgs, stack, dst
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--n-insert", type=int, default=10_000)
    ap.add_argument("--n-extract", type=int, default=4096)
    ap.add_argument("--probe", action="store_true", help="one launch of each operator, no timing (counter runs)")
    ap.add_argument("--serve-qps", type=float, default=200.0, help="request rate of the serving run")
    ap.add_argument("--serve-requests", type=int, default=400, help="requests in the serving run (0: skip)")
    ap.add_argument("--serve-only", action="store_true", help="only the serving run")
    ap.add_argument("--no-neighbours", action="store_true",
                    help="skip the CTF prologue and virtual-volume timings (not part of the step)")
    ap.add_argument("--precision", action="store_true",
                    help="FP32/FP16/BF16 volume storage for the 256^3 extraction: error against the FP64 oracle")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the multi-GPU path is not built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import noa_amd.xform as nx
    from noa_amd import workloads

    dev = torch.device("cuda:0")
    if args.serve_only:
        print(json.dumps(serving_report(torch, dev, args.serve_requests, args.serve_qps)), flush=True)
        return
    log("generating inputs ...")
    ext = workloads.extract_256(dev, b=args.n_extract)
    if args.precision:
        print(json.dumps(precision_report(ext, nx, torch)), flush=True)
        return
    ins = workloads.insert_256(dev, b=args.n_insert, keep_raw=not args.probe)
    n = 256
    out = torch.empty((args.n_extract, 1, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=dev)
    wvol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    shifts = torch.from_numpy(ext.shifts).to(dev)

    def extract():
        # extract_central_slices_3d then phase_shift_2d of the slices by their in-plane shifts, fused.
        nx.extract_central_slices_3d_shifted("hc2hc", ext.volume, None, ext.volume_shape, out, None, ext.slice_shape,
                                             None, ext.rotations, shifts=shifts)

    def insert():
        nx.insert_central_slices_3d("h2h", ins.slices, ins.weights, ins.slice_shape, vol, wvol, ins.volume_shape,
                                    None, ins.inv_rotations)

    if args.probe:
        # One launch of each operator and a device copy of known size (1 GiB read + 1 GiB written), for hardware
        # counter runs: the copy calibrates FETCH_SIZE / WRITE_SIZE against known bytes.
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).normal_()
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        extract()
        insert()
        torch.cuda.synchronize()
        log("probe done: copy 1 GiB, extract, insert")
        return

    log("warmup ...")
    for _ in range(args.warmup):
        extract()
        insert()
    torch.cuda.synchronize()

    # K steps bracketed by synchronisations; HIP events on the launch stream time each operator inside them.
    marks = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        marks[k][0].record()
        extract()
        marks[k][1].record()
        insert()
        marks[k][2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    durations = {"extract": [m[0].elapsed_time(m[1]) * 1e-3 for m in marks],
                 "insert": [m[1].elapsed_time(m[2]) * 1e-3 for m in marks]}
    ms_per_step = wall / args.steps * 1e3
    slices_per_step = args.n_extract + args.n_insert
    value = slices_per_step / (wall / args.steps)

    kernels = {}
    for name, nbytes, nsl in (("extract", algorithmic_bytes_extract(args.n_extract, n), args.n_extract),
                              ("insert", algorithmic_bytes_insert(args.n_insert, n), args.n_insert)):
        avg = statistics.mean(durations[name])
        gbs = nbytes / avg / 1e9
        kernels[name] = {"avg_ms": avg * 1e3, "algorithmic_bytes": nbytes, "achieved_GBs": gbs,
                         "frac_hbm_peak": gbs / HBM_PEAK_GBS, "slices_per_s": nsl / avg}
        log(f"{name}: {avg * 1e3:.3f} ms, {gbs:.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:.1f}% of HBM peak), "
            f"{nsl / avg / 1e6:.2f} M slices/s")
    dom = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    total_bytes = sum(k["algorithmic_bytes"] for k in kernels.values())
    total_time = sum(k["avg_ms"] for k in kernels.values()) * 1e-3

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"256^3 extract {args.n_extract} slices (hc, LINEAR) with their in-plane shifts applied by a "
                        f"fused phase_shift_2d + insert {args.n_insert} CTF-weighted slices with CTF^2 weights (h2h, "
                        "LINEAR, windowed sinc 1 voxel)",
            "volume": [256, 256, 256],
            "slices_per_step": slices_per_step,
            "combined_frac_hbm_peak": total_bytes / total_time / 1e9 / HBM_PEAK_GBS,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": {"insert": "insert_wave_kernel (+ insert_reduce_kernel, 2% of the operator)",
                       "extract": "extract_kernel"}[dom],
            "achieved": kernels[dom]["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": kernels[dom]["frac_hbm_peak"],
            "traffic": MEASURED_TRAFFIC.get((dom, args.n_insert if dom == "insert" else args.n_extract)),
        },
        "kernels": kernels,
    }
    if not args.no_neighbours:
        log("prologue / virtual volume ...")
        result["neighbours"] = neighbours_report(torch, dev, nx, ins, n)
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(ext, ins)
    if args.serve_requests > 0:
        log("serving ...")
        del ext, ins, out
        torch.cuda.empty_cache()
        result["serving"] = serving_report(torch, dev, args.serve_requests, args.serve_qps)
        log(f"serving: p50 {result['serving']['p50_ms']:.2f} ms, p99 {result['serving']['p99_ms']:.2f} ms, "
            f"{result['serving']['slices_per_s']:.0f} slices/s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
