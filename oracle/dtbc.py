"""Oracle: exact discrete transparent boundary condition (CPU, float64).

Restates src/magnelio/ports/_modal/dtbc.py:
  lambda_symbol     :91-122   outgoing root of lambda^2 - 2A lambda + 1 = 0,
                              A = 1 + (z - 2 + 1/z + q^2)/(2 r^2), |lambda| <= 1
  dtbc_kernel       :125-156  Laurent coefficients l_m of lambda(z) around
                              infinity by contour integration on |z| = rho,
                              rho = exp(4/n_kernel), FFT of 8*n_kernel points
  destagger_theta   :201-231  theta = -log(lambda)/2 at z = e^{j w dt}(1+1e-8)
  dtbc_wave_impedance :234-287  Z_TE = z0 s/rad, Z_TM = z0 rad/s
  DTBCTermination   :290-463  u^{n+1} = (2 - 2r^2 - q^2) u^n - u^{n-1}
                              + r^2 (ghost^n + u_int^n),
                              ghost = s^n + sum_{m>=1} l_m w^{n-m},
                              w^n = u^n - sum_{m>=1} l_m s^{n-m};
                              kernel length 4096, doubled and regenerated
                              whenever a step outgrows it.

Parity pins (tests/test_oracle_dtbc.py): the kernel reconstructs the
symbol on the contour (the reference's own gate, test_dtbc.py:97-104),
l_0 = 0 and l_1 = r^2 (leading Laurent term), the advance equals a
brute-force convolution, and a long explicit 1-D chain terminated by the
DTBC is transparent to better than -100 dB against an unterminated
infinite-line reference (test_dtbc.py:260-296).
"""

from __future__ import annotations

import math

import numpy as np

EDGE_OFFSET = 1e-8
N_KERNEL_INIT = 4096


def lambda_symbol(z, r: float, q: float = 0.0):
    z = np.asarray(z, dtype=complex)
    a = 1.0 + (z - 2.0 + 1.0 / z + q * q) / (2.0 * r * r)
    root = np.sqrt(a * a - 1.0)
    lm, lp = a - root, a + root
    return np.where(np.abs(lm) <= np.abs(lp), lm, lp)


def dtbc_kernel(r: float, q: float, n_kernel: int) -> np.ndarray:
    if n_kernel < 2:
        raise ValueError("n_kernel must be >= 2")
    n_fft = 8 * n_kernel
    rho = math.exp(4.0 / n_kernel)
    theta = 2.0 * np.pi * np.arange(n_fft) / n_fft
    coeff = np.fft.ifft(lambda_symbol(rho * np.exp(1j * theta), r, q))
    return np.real(coeff[:n_kernel]) * rho ** np.arange(n_kernel)


def kernel_length_for_step(n: int, init: int = N_KERNEL_INIT) -> int:
    """Kernel length in force at step n (0-based): doubled until > n."""
    k = init
    while k < n + 1:
        k *= 2
    return k


def destagger_theta(w_dt, r: float, q: float = 0.0):
    z = np.exp(1j * np.asarray(w_dt, dtype=float)) * (1.0 + EDGE_OFFSET)
    return -0.5 * np.log(lambda_symbol(z, r, q))


def dtbc_wave_impedance(w_dt, q: float, z0: float, kind: str):
    s = np.sin(np.asarray(w_dt, dtype=float) / 2.0)
    d = s * s - (q / 2.0) ** 2
    rad = np.where(d >= 0.0, np.sqrt(np.maximum(d, 0.0)) + 0.0j, -1j * np.sqrt(np.maximum(-d, 0.0)))
    if kind == "TE":
        return z0 * s / rad
    if kind == "TM":
        return z0 * rad / s
    raise ValueError(f"mode_kind must be 'TE' or 'TM', got {kind!r}")


class DTBC:
    """One boundary chain; ``advance`` mirrors DTBCTermination.advance."""

    def __init__(self, r: float, q: float = 0.0, n_kernel_init: int = N_KERNEL_INIT):
        if not (0.0 < r <= 1.0):
            raise ValueError(f"r must be in (0, 1], got {r}")
        self.r, self.q = float(r), float(q)
        self.r2, self.q2 = self.r * self.r, self.q * self.q
        self.nk = int(n_kernel_init)
        self.kernel = dtbc_kernel(self.r, self.q, self.nk)
        self.u = 0.0
        self.u_prev = 0.0
        self.n = 0
        self.w = np.zeros(self.nk)
        self.s = np.zeros(self.nk)

    def _grow(self, n: int) -> None:
        if n + 1 > self.nk:
            while self.nk < n + 1:
                self.nk *= 2
            self.kernel = dtbc_kernel(self.r, self.q, self.nk)
        if n + 1 > self.w.size:
            cap = max(2 * self.w.size, n + 1)
            self.w = np.concatenate([self.w, np.zeros(cap - self.w.size)])
            self.s = np.concatenate([self.s, np.zeros(cap - self.s.size)])

    def advance(self, u_int: float, src: float = 0.0) -> float:
        n = self.n
        self._grow(n)
        if n == 0:
            u_inc = conv_w = 0.0
        else:
            lm = self.kernel[n:0:-1]  # l_n ... l_1 against hist[0 .. n-1]
            u_inc = float(np.dot(lm, self.s[:n]))
            conv_w = float(np.dot(lm, self.w[:n]))
        ghost = src + conv_w
        self.w[n] = self.u - u_inc
        self.s[n] = src
        u_new = (2.0 - 2.0 * self.r2 - self.q2) * self.u - self.u_prev + self.r2 * (ghost + u_int)
        self.u_prev = self.u
        self.u = u_new
        self.n = n + 1
        return u_new
