"""Seeded synthetic inputs of the central-slice workloads, generated on the device.

Large arrays come from torch's device generator (Philox, seeded); poses and per-slice parameters come from
numpy's PCG64 on the host. Both are reproducible for a given seed. noa's own `noa::randomize` seeds from
std::random_device (src/noa/runtime/Random.hpp:14) and cannot be reproduced, so this is the project's RNG.

Workloads (all FP32 unless stated):
  correctness_64   64 real 64x64 N(0,1) images, rfft'd (norm "forward", noa's default, fft/Transform.hpp:13),
                   "h" layout; 64 uniform-SO(3) rotations; 64^3 volume + weights.
  extract_256      256x256x129 complex64 "hc" volume, CN(0,1) * exp(-20 |f|^2); 4096 uniform-SO(3) poses and
                   in-plane shifts U(-5, 5) px; 256x129 output slices; LINEAR, cutoff 0.5.
  insert_256       10,000 real 256x256 N(0,1) images, rfft'd, multiplied by their CTF (300 kV, Cs 2.7 mm,
                   amplitude 0.07, 1.0 A/px, defocus U(0.5, 3.0) um, astigmatism U(0, 0.1) um at U(0, pi));
                   weights CTF^2; inserted into a 256^3 volume with the inverse rotations.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from .xform import Mat33Array, euler2matrix

SEED_BASE = 0x5EED0000


def uniform_rotations(n: int, rng: np.random.Generator, dtype=np.float32) -> np.ndarray:
    """Rotations uniform on SO(3): ZYZ Euler (phi, theta, psi) with theta = arccos(U(-1, 1)), through
    noa's euler2matrix (xform/core/Euler.cpp:218-274, axes "zyz", intrinsic, right-handed)."""
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    psi = rng.uniform(0, 2 * np.pi, n)
    return euler2matrix(np.stack([phi, theta, psi], -1), axes="zyz", dtype=dtype)


def fftfreq_grid(n: int, centered: bool, device) -> torch.Tensor:
    i = torch.arange(n, device=device)
    f = (i - n // 2) if centered else torch.where(i < (n + 1) // 2, i, i - n)
    return f.to(torch.float32) / n


def relativistic_electron_wavelength(volts: float) -> float:
    """signal/core/CTF.hpp:12-25 (metres)."""
    h, c, m0, e = 6.62607015e-34, 299792458.0, 9.1093837015e-31, 1.602176634e-19
    ev = e * volts
    return h * c / math.sqrt(ev * (2 * m0 * c * c + ev))


def ctf_anisotropic(shape_hw, defocus_um, astig_um, angle, *, voltage_kv=300.0, cs_mm=2.7, amplitude=0.07,
                    pixel_size=1.0, centered=False, device="cuda") -> torch.Tensor:
    """CTFAnisotropic::value_at (signal/core/CTF.hpp:490-503) on an rfft grid, for a batch of defocus values.

    Returns (B, H, W//2+1) float32. No B-factor, no phase shift, scale 1."""
    h, w = shape_hw
    lam = relativistic_electron_wavelength(voltage_kv * 1e3) * 1e10   # angstroms
    k1 = math.pi * lam
    k2 = math.pi * 0.5 * (cs_mm * 1e7) * lam ** 3
    k3 = math.atan(amplitude / math.sqrt(1 - amplitude * amplitude))
    fy = fftfreq_grid(h, centered, device)[:, None] / pixel_size
    fx = (torch.arange(w // 2 + 1, device=device, dtype=torch.float32) / w)[None, :] / pixel_size
    phi = torch.atan2(fy, fx)
    r2 = fy * fy + fx * fx
    d = torch.as_tensor(-np.asarray(defocus_um) * 1e4, dtype=torch.float32, device=device)[:, None, None]
    a = torch.as_tensor(-np.asarray(astig_um) * 1e4, dtype=torch.float32, device=device)[:, None, None]
    ang = torch.as_tensor(np.asarray(angle), dtype=torch.float32, device=device)[:, None, None]
    defocus = d + a * torch.cos(2 * (phi[None] - ang))
    phase = k1 * r2[None] * defocus + k2 * (r2 * r2)[None] - k3
    return -torch.sin(phase)


@dataclass
class ExtractWorkload:
    volume: torch.Tensor           # (1, N, N, N//2+1) complex64, "hc"
    volume_shape: tuple
    slice_shape: tuple
    rotations: Mat33Array          # forward rotations
    poses: np.ndarray              # (B, 3, 3) float32 host copy
    shifts: np.ndarray             # (B, 2) float32, in pixels


@dataclass
class InsertWorkload:
    slices: torch.Tensor           # (B, 1, N, N//2+1) complex64, "h", CTF-multiplied
    weights: torch.Tensor          # (B, 1, N, N//2+1) float32, CTF^2
    slice_shape: tuple
    volume_shape: tuple
    inv_rotations: Mat33Array      # inverse rotations (insert_central_slices_3d takes R^T)
    poses: np.ndarray              # forward rotations, host copy
    raw: torch.Tensor | None = None       # (B, 1, N, N//2+1) complex64 spectra before the CTF (keep_raw)
    ctfs: torch.Tensor | None = None      # (B, 56) uint8: the slices' CTFAnisotropic<f32> objects (keep_raw)


def extract_256(device="cuda", n: int = 256, b: int = 4096, seed: int = SEED_BASE + 2) -> ExtractWorkload:
    g = torch.Generator(device=device).manual_seed(seed)
    shape = (1, n, n, n // 2 + 1)
    re = torch.randn(shape, generator=g, device=device)
    im = torch.randn(shape, generator=g, device=device)
    fz = fftfreq_grid(n, True, device)[:, None, None]
    fy = fftfreq_grid(n, True, device)[None, :, None]
    fx = (torch.arange(n // 2 + 1, device=device, dtype=torch.float32) / n)[None, None, :]
    env = torch.exp(-20 * (fz * fz + fy * fy + fx * fx))
    vol = torch.complex(re * env, im * env).contiguous()
    del re, im
    rng = np.random.default_rng(seed)
    poses = uniform_rotations(b, rng)
    shifts = rng.uniform(-5, 5, (b, 2)).astype(np.float32)
    return ExtractWorkload(vol, (1, n, n, n), (b, 1, n, n), Mat33Array(poses, device), poses, shifts)


def insert_256(device="cuda", n: int = 256, b: int = 10_000, seed: int = SEED_BASE + 3,
               chunk: int = 1000, keep_raw: bool = False) -> InsertWorkload:
    g = torch.Generator(device=device).manual_seed(seed)
    rng = np.random.default_rng(seed)
    poses = uniform_rotations(b, rng)
    defocus = rng.uniform(0.5, 3.0, b)
    astig = rng.uniform(0.0, 0.1, b)
    angle = rng.uniform(0.0, np.pi, b)
    slices = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.complex64, device=device)
    weights = torch.empty((b, 1, n, n // 2 + 1), dtype=torch.float32, device=device)
    raw = torch.empty_like(slices) if keep_raw else None
    for s in range(0, b, chunk):
        e = min(b, s + chunk)
        img = torch.randn((e - s, n, n), generator=g, device=device)
        spec = torch.fft.rfft2(img, norm="forward")
        ctf = ctf_anisotropic((n, n), defocus[s:e], astig[s:e], angle[s:e], device=device)
        slices[s:e, 0] = spec * ctf
        weights[s:e, 0] = ctf * ctf
        if keep_raw:
            raw[s:e, 0] = spec
        del img, spec, ctf
    inv = np.ascontiguousarray(np.transpose(poses, (0, 2, 1)))
    ctfs = None
    if keep_raw:
        from .signal import CTFAnisotropic, ctf_array
        ctfs = ctf_array([CTFAnisotropic((1.0, 1.0), (defocus[i], astig[i], angle[i]), 300.0, 0.07, 2.7)
                          for i in range(b)], device).reshape(b, 56)
    return InsertWorkload(slices, weights, (b, 1, n, n), (1, n, n, n), Mat33Array(inv, device), poses, raw, ctfs)


def correctness_64(device="cuda", n: int = 64, b: int = 64, seed: int = SEED_BASE + 1):
    g = torch.Generator(device=device).manual_seed(seed)
    img = torch.randn((b, n, n), generator=g, device=device)
    slices = torch.fft.rfft2(img, norm="forward").reshape(b, 1, n, n // 2 + 1).contiguous()
    rng = np.random.default_rng(seed)
    poses = uniform_rotations(b, rng)
    return slices, poses


SERVING_BOXES = (64, 128, 192, 256, 384)
SERVING_P = (0.3, 0.3, 0.2, 0.15, 0.05)


def serving_volumes(device="cuda", per_box: int = 4, seed: int = SEED_BASE + 5) -> dict:
    """4 resident complex64 "hc" volumes per box size, CN(0,1) x exp(-20 |f|^2) as the extraction workload."""
    g = torch.Generator(device=device).manual_seed(seed)
    vols = {}
    for n in SERVING_BOXES:
        fz = fftfreq_grid(n, True, device)[:, None, None]
        fy = fftfreq_grid(n, True, device)[None, :, None]
        fx = (torch.arange(n // 2 + 1, device=device, dtype=torch.float32) / n)[None, None, :]
        env = torch.exp(-20 * (fz * fz + fy * fy + fx * fx))
        vols[n] = []
        for _ in range(per_box):
            shape = (1, n, n, n // 2 + 1)
            re = torch.randn(shape, generator=g, device=device)
            im = torch.randn(shape, generator=g, device=device)
            vols[n].append(torch.complex(re * env, im * env).contiguous())
    return vols


def serving_requests(device="cuda", n_requests: int = 300, qps: float = 200.0, seed: int = SEED_BASE + 5,
                     max_k: int = 512):
    """A Poisson request stream at `qps`: box from {64, 128, 192, 256, 384} with p {.3, .3, .2, .15, .05}; half
    extract K ~ U{1..max_k} slices at uniform poses from one of the 4 resident volumes of that box, half insert K
    CTF-weighted slices (as the insertion workload, at that box) into a per-request accumulator. Inputs are
    generated on the device before the stream starts."""
    from .serve import Request
    g = torch.Generator(device=device).manual_seed(seed + 1)
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n_requests))
    reqs = []
    for i in range(n_requests):
        box = int(rng.choice(SERVING_BOXES, p=SERVING_P))
        k = int(rng.integers(1, max_k + 1))
        poses = uniform_rotations(k, rng)
        if rng.uniform() < 0.5:
            reqs.append(Request(rid=i, kind="extract", box=box, volume=int(rng.integers(0, 4)),
                                rotations=torch.from_numpy(poses).to(device), arrival=float(arrivals[i])))
        else:
            img = torch.randn((k, box, box), generator=g, device=device)
            spec = torch.fft.rfft2(img, norm="forward")
            ctf = ctf_anisotropic((box, box), rng.uniform(0.5, 3.0, k), rng.uniform(0.0, 0.1, k),
                                  rng.uniform(0.0, np.pi, k), device=device)
            inv = np.ascontiguousarray(np.transpose(poses, (0, 2, 1)))
            reqs.append(Request(rid=i, kind="insert", box=box, rotations=torch.from_numpy(inv).to(device),
                                slices=(spec * ctf).reshape(k, 1, box, box // 2 + 1).contiguous(),
                                weights=(ctf * ctf).reshape(k, 1, box, box // 2 + 1).contiguous(),
                                arrival=float(arrivals[i])))
            del img, spec, ctf
    return reqs
