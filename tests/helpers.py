"""Shared test helpers: comparisons, seeded inputs and layout remaps (numpy)."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def allclose_abs_safe(a, b, eps: float) -> bool:
    """tests/Utils.hpp:330-357: per real component, |a-b| <= eps or |a-b| <= max(|a|,|b|)*eps; non-finite fails."""
    a = np.asarray(a)
    b = np.asarray(b)
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        a = np.stack([np.real(a), np.imag(a)])
        b = np.stack([np.real(b), np.imag(b)])
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    if not (np.isfinite(a).all() and np.isfinite(b).all()):
        return False
    d = np.abs(a - b)
    return bool(np.all((d <= eps) | (d <= np.maximum(np.abs(a), np.abs(b)) * eps)))


def random_rotations(n: int, rng: np.random.Generator, dtype=np.float32) -> np.ndarray:
    """Rotations uniform on SO(3): ZYZ Euler angles phi, psi ~ U(0, 2pi), theta = arccos(U(-1, 1))."""
    from noa_amd.xform import euler2matrix
    phi = rng.uniform(0, 2 * np.pi, n)
    theta = np.arccos(rng.uniform(-1, 1, n))
    psi = rng.uniform(0, 2 * np.pi, n)
    return euler2matrix(np.stack([phi, theta, psi], -1), axes="zyz", dtype=dtype)


def random_spectrum(shape, rng: np.random.Generator, dtype=np.complex64) -> np.ndarray:
    if np.dtype(dtype).kind == "c":
        return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    return rng.standard_normal(shape).astype(dtype)


def h2hc(x: np.ndarray, ndim: int) -> np.ndarray:
    """noa::fft::remap("h2hc") of an rfft array: fftshift along the `ndim - 1` leading spatial axes."""
    axes = tuple(range(-ndim, -1))
    return np.fft.fftshift(x, axes=axes)


def hc2h(x: np.ndarray, ndim: int) -> np.ndarray:
    axes = tuple(range(-ndim, -1))
    return np.fft.ifftshift(x, axes=axes)


def freq_grid(n: int, centered: bool) -> np.ndarray:
    """Frequencies of the indices of a full dimension (numpy's fftfreq, an independent routine)."""
    f = np.fft.fftfreq(n, 1.0 / n).round().astype(np.int64)
    return np.fft.fftshift(f) if centered else f


def mat33_blocks(r: np.ndarray) -> np.ndarray:
    """(B, 3, 3) -> noa's Mat33<f32> memory layout (48-byte stride)."""
    out = np.zeros((r.shape[0], 12 if r.dtype == np.float32 else 10), dtype=r.dtype)
    out[:, :9] = r.reshape(r.shape[0], 9)
    return out
