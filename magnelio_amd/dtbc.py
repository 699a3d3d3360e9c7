"""Host-side (float64) math of the exact discrete transparent boundary.

The per-step recursion runs on the device (csrc/mgh_ports.hip); the host
builds the convolution kernels it consumes and the closed forms the
S-parameter post-processing needs.

Reference: src/magnelio/ports/_modal/dtbc.py
  lambda_symbol :91-122, dtbc_kernel :125-156, destagger_theta :201-231,
  dtbc_wave_impedance :234-287, kernel schedule (initial length 4096,
  doubled and regenerated when a step outgrows it) :345-363.

Kernel: l_m are the Laurent coefficients of lambda(z) around infinity,
obtained by contour integration on |z| = rho = exp(4/n) with an 8n-point
FFT, times rho^m.  Because the contour radius depends on the kernel length,
each regenerated kernel differs from its predecessor in the last bits; the
device therefore receives every kernel version the run will use and
switches at the same steps the reference regenerates.
"""

from __future__ import annotations

import math

import numpy as np

EDGE_OFFSET = 1e-8
N_KERNEL_INIT = 4096


def lambda_symbol(z, r: float, q: float = 0.0) -> np.ndarray:
    """Outgoing root |lambda| <= 1 of lambda^2 - 2 A lambda + 1 = 0."""
    z = np.asarray(z, dtype=complex)
    a = 1.0 + (z - 2.0 + 1.0 / z + q * q) / (2.0 * r * r)
    root = np.sqrt(a * a - 1.0)
    lam_m = a - root
    lam_p = a + root
    return np.where(np.abs(lam_m) <= np.abs(lam_p), lam_m, lam_p)


def dtbc_kernel(r: float, q: float, n_kernel: int) -> np.ndarray:
    """Convolution kernel l_0 .. l_{n-1} (l_0 = 0 analytically)."""
    if n_kernel < 2:
        raise ValueError("n_kernel must be >= 2")
    n_fft = 8 * n_kernel
    rho = math.exp(4.0 / n_kernel)
    theta = 2.0 * np.pi * np.arange(n_fft) / n_fft
    coeff = np.fft.ifft(lambda_symbol(rho * np.exp(1j * theta), r, q))
    return np.real(coeff[:n_kernel]) * rho ** np.arange(n_kernel)


def kernel_versions(n_steps: int, init: int = N_KERNEL_INIT) -> list[int]:
    """Kernel lengths in force over steps 0 .. n_steps-1 (init, 2 init, ...)."""
    out = [init]
    while out[-1] < n_steps:
        out.append(out[-1] * 2)
    return out


def packed_kernels(r: float, q: float, n_steps: int, init: int = N_KERNEL_INIT) -> np.ndarray:
    """All kernel versions of one channel concatenated: version v (length
    init*2^v) starts at offset init*(2^v - 1)."""
    return np.concatenate([dtbc_kernel(r, q, n) for n in kernel_versions(n_steps, init)])


def destagger_theta(w_dt, r: float, q: float = 0.0) -> np.ndarray:
    """theta = -log(lambda)/2 at z = e^{j w dt} (1 + 1e-8)."""
    z = np.exp(1j * np.asarray(w_dt, dtype=float)) * (1.0 + EDGE_OFFSET)
    return -0.5 * np.log(lambda_symbol(z, r, q))


def dtbc_wave_impedance(w_dt, q: float, z0: float, mode_kind: str) -> np.ndarray:
    """Exact discrete wave impedance Z_TE = z0 s/rad, Z_TM = z0 rad/s."""
    s = np.sin(np.asarray(w_dt, dtype=float) / 2.0)
    d = s * s - (q / 2.0) ** 2
    rad = np.where(d >= 0.0, np.sqrt(np.maximum(d, 0.0)) + 0.0j, -1j * np.sqrt(np.maximum(-d, 0.0)))
    if mode_kind == "TE":
        return z0 * s / rad
    if mode_kind == "TM":
        return z0 * rad / s
    raise ValueError(f"mode_kind must be 'TE' or 'TM', got {mode_kind!r}")
