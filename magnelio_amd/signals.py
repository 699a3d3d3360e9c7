"""Excitation waveforms, sampled signals and power-wave S-parameters (host, float64).

Reference counterparts:

* ``gaussian`` / ``modulated_gaussian``   src/magnelio/signals/waveforms.py:28-81
  (sigma = 2/(pi B), t0 = 4/B, carrier at the band centre)
* ``resolve_excitation``                   src/magnelio/analysis/scattering_td.py:2322-2366
  (lower band edge max(f_cutoff, f_min); modulated Gaussian above cut-off)
* ``Signal1D.at_frequencies``              src/magnelio/signals/signal_1d.py:60-118
  (direct DFT sum_n x_n exp(-2 pi j f n dt) when Nf*N <= 1e8)
* ``spectral_power_waves``                 src/magnelio/post/modal_sparameters.py:118-242
* ``compute_s_parameters``                 src/magnelio/post/modal_sparameters.py:335-557
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .dtbc import destagger_theta, dtbc_wave_impedance


def gaussian(t, f_max: float):
    sigma = 2.0 / (math.pi * f_max)
    t0 = 4.0 / f_max
    x = (np.asarray(t, dtype=float) - t0) / sigma
    return np.exp(-x * x)


def modulated_gaussian(t, f_max: float, f_min: float):
    bw = f_max - f_min
    sigma = 2.0 / (math.pi * bw)
    t0 = 4.0 / bw
    t = np.asarray(t, dtype=float)
    x = (t - t0) / sigma
    fc = 0.5 * (f_min + f_max)
    return np.exp(-x * x) * np.cos(2.0 * math.pi * fc * (t - t0))


def _scalar_gauss(t: float, f_max: float) -> float:
    sigma = 2.0 / (math.pi * f_max)
    t0 = 4.0 / f_max
    x = (t - t0) / sigma
    return math.exp(-x * x)


def _scalar_mod_gauss(t: float, f_max: float, f_min: float) -> float:
    bw = f_max - f_min
    sigma = 2.0 / (math.pi * bw)
    t0 = 4.0 / bw
    x = (t - t0) / sigma
    fc = 0.5 * (f_min + f_max)
    return math.exp(-x * x) * math.cos(2.0 * math.pi * fc * (t - t0))


@dataclass(frozen=True)
class Excitation:
    """Band-limited port excitation (factory.py ExcitationSpec :142-178)."""

    f_min: float
    f_max: float
    mode_index: int = 0
    waveform: str = "modulated_gaussian"

    def __call__(self, t: float) -> float:
        if self.waveform == "modulated_gaussian":
            return _scalar_mod_gauss(float(t), self.f_max, self.f_min)
        if self.waveform == "gaussian":
            return _scalar_gauss(float(t), self.f_max)
        raise ValueError(f"unknown waveform {self.waveform!r}")

    def sample(self, t: np.ndarray) -> np.ndarray:
        """Scalar-identical samples (the scalar math path, element by element)."""
        return np.array([self(float(v)) for v in np.asarray(t, dtype=float)])


def resolve_excitation(omega_c: float, f_min: float, f_max: float, mode_index: int = 0) -> Excitation:
    f_cut = omega_c / (2.0 * math.pi)
    edge = max(f_cut, f_min)
    if edge >= f_max:
        raise ValueError(f"f_max = {f_max:.4g} Hz does not exceed the lower band edge {edge:.4g} Hz")
    wf = "gaussian" if edge <= 0.0 else "modulated_gaussian"
    return Excitation(edge, f_max, mode_index, wf)


def estimate_steps(grid, excitation: Excitation, dt: float, n_traversals: int = 25) -> int:
    """2 t0 + 25 diagonal transits at c0/2 (scattering_td.py:2601-2643)."""
    from .mesh import C0  # noqa: PLC0415

    bw = max(excitation.f_max - max(excitation.f_min, 0.0), excitation.f_max)
    t0 = 4.0 / bw
    lx, ly, lz = (float(a[-1] - a[0]) for a in (grid.x, grid.y, grid.z))
    diag = math.sqrt(lx * lx + ly * ly + lz * lz)
    t_total = 2.0 * t0 + n_traversals * (diag / (0.5 * C0) if diag > 0 else 0.0)
    return int(math.ceil(t_total / dt))


def dft(values: np.ndarray, f: np.ndarray, dt: float) -> np.ndarray:
    """sum_n x_n exp(-2 pi j f n dt) (no dt factor), direct evaluation."""
    values = np.asarray(values, dtype=float)
    f = np.asarray(f, dtype=float).ravel()
    n = np.arange(values.size)
    if f.size * values.size <= int(1e8):
        return np.exp(-2j * math.pi * np.outer(f, n) * dt) @ values
    out = np.empty(f.size, dtype=complex)
    for lo in range(0, f.size, 64):
        fs = f[lo:lo + 64]
        out[lo:lo + 64] = np.exp(-2j * math.pi * np.outer(fs, n) * dt) @ values
    return out


def power_waves(V_f, I_f, omega, dt, mode, normal_dx=None, line_params=None):
    """a, b of one channel (spectral_power_waves, unguarded S-parameter convention)."""
    omega = np.asarray(omega, dtype=float)
    if line_params is not None and len(line_params) > 2 and line_params[2] is not None:
        Z = dtbc_wave_impedance(omega * dt, line_params[1], line_params[2], mode.mode_type)
    else:
        Z = np.array([mode.z_modal(float(w)) for w in omega], dtype=complex)
    with np.errstate(divide="ignore", invalid="ignore"):
        sz = np.sqrt(Z)
        if normal_dx is not None:
            if line_params is not None:
                theta = destagger_theta(omega * dt, line_params[0], line_params[1])
            else:
                theta = 0.5 * float(normal_dx) * np.array([mode.gamma(float(w)) for w in omega])
            ep, em = np.exp(+theta), np.exp(-theta)
            den = ep + em
            a = (V_f / sz * ep + sz * I_f) / den
            b = (V_f / sz * em - sz * I_f) / den
        else:
            a = (V_f / sz + sz * I_f) / 2.0
            b = (V_f / sz - sz * I_f) / 2.0
    return a, b


def s_parameters(signals: dict, port_modes: dict, excited: tuple, f_axis: np.ndarray, dt: float,
                 *, port_normal_dx: dict | None = None, port_line_params: dict | None = None,
                 lumped_z0: dict | None = None, a_threshold: float = 1e-12) -> dict:
    """S(f) per recorded channel for one excitation (compute_s_parameters).

    ``signals``: {(port, mode): (V array, I array)}; modal I carries the
    temporal Yee half-step exp(+j w dt/2) correction here.  Ports named in
    ``lumped_z0`` ({port: Z0}) are discrete ports: their V and I are
    co-temporal (the Thevenin current of the same step), so the split is
    a = (V + Z0 I) / (2 sqrt(Z0)), b = (V - Z0 I) / (2 sqrt(Z0)) without the
    half-step factor (modal_sparameters.py, test_lumped_port.py:241-285).
    """
    f_axis = np.asarray(f_axis, dtype=float)
    if np.any(f_axis <= 0):
        raise ValueError("f_axis must contain only positive frequencies")
    if excited not in signals:
        raise KeyError(f"excited channel {excited!r} not recorded")
    omega = 2.0 * math.pi * f_axis
    phase = np.exp(+1j * omega * dt / 2.0)
    a_ch, b_ch = {}, {}
    for key, (V, I) in signals.items():
        port, m = key
        if lumped_z0 is not None and port in lumped_z0:
            z0 = float(lumped_z0[port])
            V_f = dft(V, f_axis, dt)
            I_f = dft(I, f_axis, dt)
            a_ch[key] = (V_f + z0 * I_f) / (2.0 * math.sqrt(z0))
            b_ch[key] = (V_f - z0 * I_f) / (2.0 * math.sqrt(z0))
            continue
        mode = port_modes[port][m]
        V_f = dft(V, f_axis, dt)
        I_f = dft(I, f_axis, dt) * phase
        lp = port_line_params.get(key) if port_line_params else None
        ndx = port_normal_dx.get(port) if port_normal_dx else None
        a_ch[key], b_ch[key] = power_waves(V_f, I_f, omega, dt, mode, ndx, lp)
    a_exc = a_ch[excited]
    peak = float(np.max(np.abs(a_exc)))
    if peak == 0.0:
        return {k: np.full(f_axis.size, np.nan, dtype=complex) for k in signals}
    valid = np.abs(a_exc) >= a_threshold * peak
    safe = np.where(valid, a_exc, 1.0 + 0j)
    return {k: np.where(valid, b / safe, np.nan + 1j * np.nan) for k, b in b_ch.items()}
