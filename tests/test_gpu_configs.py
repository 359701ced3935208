"""Each benchmark workload at its stated size on the GPU, checked against the oracle.

  64     64 slices of 64^2 (rfft'd N(0,1) images, "h"), uniform poses: rasterisation into a 64^3 volume + weights
         (h2h, cutoff 0.5), then extraction of the same poses (h2h, LINEAR).
  256x   256^3 "hc" volume CN(0,1) exp(-20 |f|^2), 4096 uniform poses with in-plane shifts U(-5, 5) px: extraction,
         plain and with the fused phase shift, compared with the oracle on 64 poses spread over the batch.
  256i   10,000 CTF-weighted slices of 256^2 with CTF^2 weights inserted into a 256^3 volume (h2h, LINEAR, default
         windowed sinc): the GPU volume against noa's per-voxel sums, evaluated by the oracle on a voxel subset that
         holds the voxels around the origin (met by most slices), the cutoff shell and uniform samples.
  256p   the 256^3 volume stored as FP16 and BF16 pairs: error of the extraction against the FP64 oracle.

Bars: extraction bit-exact (the fused shift: rel-L2 3e-7, sin/cos of two libraries); rasterisation rel-L2 1e-6
(float atomics reorder the sums); insertion rel-L2 1e-5 on the subset and deterministic run to run.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import oracle as O
from helpers import rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda"
THREADS = 16


def nx():
    import noa_amd.xform as m
    return m


def wl():
    from noa_amd import workloads
    return workloads


def test_correctness_64_rasterize_then_extract():
    print("case workload 64: rasterize h2h 64 slices of 64^2, extract h2h", flush=True)
    n, b = 64, 64
    slices, poses = wl().correctness_64(DEV)
    sl = slices.cpu().numpy()
    rot = torch.from_numpy(poses).to(DEV)

    vol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    wvol = torch.zeros((1, n, n, n // 2 + 1), dtype=torch.float32, device=DEV)
    nx().rasterize_central_slices_3d("h2h", slices, None, (b, 1, n, n), vol, wvol, (1, n, n, n), None, rot,
                                     nx().RasterizeCentralSlicesOptions(fftfreq_cutoff=0.5))
    ovol = np.zeros((1, n, n, n // 2 + 1), np.complex64)
    owvol = np.zeros((1, n, n, n // 2 + 1), np.float32)
    O.rasterize_central_slices_3d("h2h", sl, None, (b, 1, n, n), ovol, owvol, (1, n, n, n), None, poses,
                                  fftfreq_cutoff=0.5)
    torch.cuda.synchronize()
    gv, gw = vol.cpu().numpy(), wvol.cpu().numpy()
    assert rel_l2(gv, ovol) < 1e-6
    assert rel_l2(gw, owvol) < 1e-6
    assert owvol.sum() > 0

    # Extraction of the oracle's volume: bit-exact.
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    nx().extract_central_slices_3d("h2h", torch.from_numpy(ovol).to(DEV), None, (1, n, n, n), out, None,
                                   (b, 1, n, n), None, rot)
    want = np.zeros((b, 1, n, n // 2 + 1), np.complex64)
    O.extract_central_slices_3d("h2h", ovol, None, (1, n, n, n), want, None, (b, 1, n, n), None, poses)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)

    # The whole chain on the GPU against the oracle's chain.
    nx().extract_central_slices_3d("h2h", vol, None, (1, n, n, n), out, None, (b, 1, n, n), None, rot)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), want) < 1e-6


@pytest.fixture(scope="module")
def extract_workload():
    print("case workload 256x: generating the 256^3 volume and 4096 poses", flush=True)
    return wl().extract_256(DEV)


def test_extract_256_4096_poses(extract_workload):
    print("case workload 256x: extraction of 4096 slices, 64 checked", flush=True)
    ext = extract_workload
    n, b = 256, 4096
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    nx().extract_central_slices_3d("hc2hc", ext.volume, None, ext.volume_shape, out, None, ext.slice_shape, None,
                                   ext.rotations)
    shifted = torch.empty_like(out)
    nx().extract_central_slices_3d_shifted("hc2hc", ext.volume, None, ext.volume_shape, shifted, None,
                                           ext.slice_shape, None, ext.rotations,
                                           shifts=torch.from_numpy(ext.shifts).to(DEV))
    pick = np.arange(0, b, b // 64)
    vol = ext.volume.cpu().numpy()
    want = np.zeros((len(pick), 1, n, n // 2 + 1), np.complex64)
    O.extract_central_slices_3d("hc2hc", vol, None, (1, n, n, n), want, None, (len(pick), 1, n, n), None,
                                np.ascontiguousarray(ext.poses[pick]), n_threads=THREADS, force_threads=True)
    torch.cuda.synchronize()
    got = out.cpu().numpy()[pick]
    np.testing.assert_array_equal(got, want)
    # Every slice is written (overwrite semantics, zeros beyond the cutoff): no slice is left all zero.
    assert bool((out.abs().amax(dim=(1, 2, 3)) > 0).all())

    wsh = want.copy()
    O.phase_shift_2d("hc2hc", wsh, wsh, (len(pick), 1, n, n), np.ascontiguousarray(ext.shifts[pick]))
    assert rel_l2(shifted.cpu().numpy()[pick], wsh) < 3e-7


def test_extract_256_half_storage_against_fp64(extract_workload):
    """The volume stored as FP16 and BF16 pairs (FP32 math): rel-L2 against the FP64 oracle on 128 of the 4096
    poses, off the cutoff circle (pixels within 1e-3 of |f| = 0.5 are the value or 0 depending on the coordinates'
    rounding). Bounds: FP32 1e-5, FP16 5e-4, BF16 4e-3; the errors are ordered FP32 < FP16 < BF16."""
    print("case workload 256p: FP32/FP16/BF16 storage vs FP64", flush=True)
    ext = extract_workload
    n, b, k = 256, 4096, 128
    vol64 = ext.volume.cpu().numpy().astype(np.complex128)
    ref = np.zeros((k, 1, n, n // 2 + 1), np.complex128)
    O.extract_central_slices_3d("hc2hc", vol64, None, (1, n, n, n), ref, None, (k, 1, n, n), None,
                                ext.poses[:k].astype(np.float64), n_threads=THREADS, force_threads=True)
    fy = (np.arange(n) - n // 2)[:, None] / n
    fx = np.arange(n // 2 + 1)[None, :] / n
    keep = np.broadcast_to(np.abs(np.sqrt(fy ** 2 + fx ** 2) - 0.5) > 1e-3, ref.shape)
    out = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=DEV)
    errs = {}
    for name, v in (("fp32", ext.volume), ("fp16", ext.volume.to(torch.complex32)),
                    ("bf16", nx().ComplexBF16.from_complex(ext.volume))):
        nx().extract_central_slices_3d("hc2hc", v, None, ext.volume_shape, out, None, ext.slice_shape, None,
                                       ext.rotations)
        torch.cuda.synchronize()
        got = out[:k].cpu().numpy().astype(np.complex128)
        errs[name] = rel_l2(got[keep], ref[keep])
        print(f"  {name}: rel-L2 {errs[name]:.3g} (all pixels {rel_l2(got, ref):.3g})", flush=True)
    assert errs["fp32"] < 1e-5
    assert errs["fp16"] < 5e-4
    assert errs["bf16"] < 4e-3
    assert errs["fp32"] < errs["fp16"] < errs["bf16"]


def _voxel_subset(n, rng):
    """(z, y, x) voxels of the stored half of an n^3 h2h volume: every voxel within 6 samples of the origin, 4096
    voxels in the shell 0.49 <= |f| <= 0.5, 8192 uniform samples, and two whole 8x8x16 blocks at mid radius."""
    fr = np.fft.fftfreq(n, 1.0 / n).round().astype(np.int64)          # index -> frequency, non-centred
    idx = np.arange(n)
    zz, yy, xx = np.meshgrid(idx, idx, np.arange(n // 2 + 1), indexing="ij")
    rad = np.sqrt(fr[zz] ** 2 + fr[yy] ** 2 + xx ** 2.0)
    allv = np.stack([zz.ravel(), yy.ravel(), xx.ravel()], -1)
    r = rad.ravel()
    near = allv[r <= 6]
    shell_all = allv[(r >= 0.49 * n) & (r <= 0.5 * n)]
    shell = shell_all[rng.choice(len(shell_all), 4096, replace=False)]
    uni = allv[rng.choice(len(allv), 8192, replace=False)]
    blocks = []
    for z0, y0, x0 in ((40, 200, 32), (232, 24, 48)):
        bz, by, bx = np.meshgrid(np.arange(z0, z0 + 8), np.arange(y0, y0 + 8), np.arange(x0, x0 + 16), indexing="ij")
        blocks.append(np.stack([bz.ravel(), by.ravel(), bx.ravel()], -1))
    return np.unique(np.concatenate([near, shell, uni, *blocks]), axis=0)


def test_insert_256_10000_ctf_weighted_slices():
    print("case workload 256i: insertion of 10,000 CTF-weighted slices into 256^3", flush=True)
    ins = wl().insert_256(DEV)
    n = 256
    shape = (1, n, n, n // 2 + 1)

    def run():
        vol = torch.zeros(shape, dtype=torch.complex64, device=DEV)
        wvol = torch.zeros(shape, dtype=torch.float32, device=DEV)
        nx().insert_central_slices_3d("h2h", ins.slices, ins.weights, ins.slice_shape, vol, wvol,
                                      ins.volume_shape, None, ins.inv_rotations)
        torch.cuda.synchronize()
        return vol.cpu().numpy()[0], wvol.cpu().numpy()[0]

    gv, gw = run()
    gv2, gw2 = run()
    np.testing.assert_array_equal(gv, gv2)          # deterministic
    np.testing.assert_array_equal(gw, gw2)
    del gv2, gw2

    vox = _voxel_subset(n, np.random.default_rng(41))
    print(f"  oracle on {len(vox)} voxels x 10000 slices", flush=True)
    sl = ins.slices.cpu().numpy()
    w = ins.weights.cpu().numpy()
    inv = np.ascontiguousarray(np.transpose(ins.poses, (0, 2, 1)))
    want_v, want_w, inside = O.insert_central_slices_3d_voxels("h2h", sl, w, ins.slice_shape, ins.volume_shape, inv,
                                                              vox, n_threads=THREADS)
    del sl, w
    got_v = gv[vox[:, 0], vox[:, 1], vox[:, 2]]
    got_w = gw[vox[:, 0], vox[:, 1], vox[:, 2]]
    # Beyond the cutoff noa leaves the voxel untouched: exactly zero in a zeroed volume.
    assert not np.any(got_v[~inside]) and not np.any(got_w[~inside])
    e_v = rel_l2(got_v[inside], want_v[inside])
    e_w = rel_l2(got_w[inside], want_w[inside])
    print(f"  rel-L2 values {e_v:.3g}, weights {e_w:.3g} over {int(inside.sum())} voxels", flush=True)
    assert e_v < 1e-5
    assert e_w < 1e-5
    # The origin neighbourhood meets most slices: a dropped slice segment there shows in the weights. Its voxels sum
    # ~10,000 terms, where the oracle's own FP32 running sum (noa's order) drifts by ~2e-5: both are compared with the
    # same sums in FP64.
    fr = np.fft.fftfreq(n, 1.0 / n).round()
    near = (np.abs(fr[vox[:, 0]]) + np.abs(fr[vox[:, 1]]) + vox[:, 2] <= 3) & inside
    ref_v, ref_w, _ = O.insert_central_slices_3d_voxels(
        "h2h", ins.slices.cpu().numpy().astype(np.complex128), ins.weights.cpu().numpy().astype(np.float64),
        ins.slice_shape, ins.volume_shape, inv.astype(np.float64), vox[near], n_threads=THREADS)
    e_gpu = rel_l2(got_w[near], ref_w)
    e_oracle = rel_l2(want_w[near], ref_w)
    print(f"  near the origin ({int(near.sum())} voxels), weights vs FP64: GPU {e_gpu:.3g}, FP32 oracle {e_oracle:.3g}; "
          f"values vs FP64: GPU {rel_l2(got_v[near], ref_v):.3g}", flush=True)
    assert e_gpu < 1e-5
    assert rel_l2(got_v[near], ref_v) < 1e-5
    assert ref_w.min() > 0


def test_insert_256_tilt_series_default_and_pose_plans(monkeypatch):
    """4,000 CTF-weighted slices of a +-60 degree single-axis tilt series into 256^3 (orientations far from uniform):
    the default plan (work items of at most 1,024 slices) against the oracle's per-voxel gather on a voxel subset
    (rel-L2 1e-5), and against the pose-adaptive plan over the whole volume (rel-L2 1e-6: the segments regroup the
    per-voxel sums)."""
    print("case workload 256i tilt series: 4,000 slices, default and pose-adaptive plans", flush=True)
    ins = wl().insert_256(DEV)
    n, b = 256, 4000
    rng = np.random.default_rng(43)
    ang = np.stack([np.zeros(b), np.deg2rad(rng.uniform(-60, 60, b)), rng.uniform(0, 2 * np.pi, b)], -1)
    fwd = nx().euler2matrix(ang, axes="zyz", dtype=np.float32)
    inv = np.ascontiguousarray(np.transpose(fwd, (0, 2, 1)))
    slices, weights = ins.slices[:b].contiguous(), ins.weights[:b].contiguous()
    shape = (1, n, n, n // 2 + 1)

    def run(order):
        monkeypatch.setenv("NOA_HIP_INSERT_ORDER", order)
        vol = torch.zeros(shape, dtype=torch.complex64, device=DEV)
        wvol = torch.zeros(shape, dtype=torch.float32, device=DEV)
        nx().insert_central_slices_3d("h2h", slices, weights, (b, 1, n, n), vol, wvol, (1, n, n, n), None,
                                      torch.from_numpy(inv).to(DEV))
        torch.cuda.synchronize()
        return vol.cpu().numpy()[0], wvol.cpu().numpy()[0]

    gv, gw = run("0")
    pv, pw = run("3")
    assert rel_l2(gv, pv) < 1e-6 and rel_l2(gw, pw) < 1e-6
    vox = _voxel_subset(n, np.random.default_rng(44))
    want_v, want_w, inside = O.insert_central_slices_3d_voxels("h2h", slices.cpu().numpy(), weights.cpu().numpy(),
                                                              (b, 1, n, n), (1, n, n, n), inv, vox, n_threads=THREADS)
    got_v = gv[vox[:, 0], vox[:, 1], vox[:, 2]]
    got_w = gw[vox[:, 0], vox[:, 1], vox[:, 2]]
    assert not np.any(got_v[~inside]) and not np.any(got_w[~inside])
    e_v, e_w = rel_l2(got_v[inside], want_v[inside]), rel_l2(got_w[inside], want_w[inside])
    print(f"  rel-L2 values {e_v:.3g}, weights {e_w:.3g} over {int(inside.sum())} voxels", flush=True)
    assert e_v < 1e-5 and e_w < 1e-5
