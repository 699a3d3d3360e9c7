"""Oracle: port-signal DFT and power-wave S-parameters (CPU, float64).

Restates src/magnelio/signals/signal_1d.py:60-118 (direct DFT
sum_n x_n e^{-2 pi j f n dt}, no dt factor) and
src/magnelio/post/modal_sparameters.py:118-242, :335-557: I is shifted by
the temporal half step e^{+j w dt/2}; certified DTBC channels use the exact
discrete wave impedance (TE: z0 s/rad, TM: z0 rad/s) and the discrete
half-cell factor lambda^{1/2} for the two-plane de-stagger; channels on Mur
the continuum wave impedance (TE j w mu0 / gamma, TM gamma / (j w eps)) and
the factor e^{gamma dx_n / 2} (mode.py:170-230); S = b / a_excited with |a| < 1e-12 peak -> NaN.

Pinned by tests/test_oracle_sparams.py: the DFT of a sine and of DC against
closed forms, and a synthetic discrete travelling wave whose a/b split must
give |b| = 0 (reflection-free) and a equal to the launched amplitude.
"""

from __future__ import annotations

import math

import numpy as np

from .dtbc import destagger_theta, dtbc_wave_impedance


def dft(x, f, dt):
    n = np.arange(len(x))
    return np.exp(-2j * math.pi * np.outer(np.asarray(f, float), n) * dt) @ np.asarray(x, float)


def waves(V, I, f, dt, line, mode=None, normal_dx=None):
    """(a, b) at the port plane of a TE or TM channel: certified (line =
    (r, q, z0)) or on Mur (line None, mode = (omega_c, eps_r[, kind]),
    normal_dx the port cell)."""
    from .ports import z_mode  # noqa: PLC0415

    kind = mode[2] if mode is not None and len(mode) > 2 else "TE"
    w = 2 * math.pi * np.asarray(f, float)
    Vf = dft(V, f, dt)
    If = dft(I, f, dt) * np.exp(1j * w * dt / 2)
    if line is not None:
        r, q, z0 = line
        Z = dtbc_wave_impedance(w * dt, q, z0, kind)
        th = destagger_theta(w * dt, r, q)
    else:
        zg = [z_mode(kind, float(om), mode[0], mode[1]) for om in w]
        Z = np.array([a for a, _ in zg], complex)
        th = 0.5 * float(normal_dx) * np.array([g for _, g in zg], complex)
    sz = np.sqrt(Z)
    ep, em = np.exp(th), np.exp(-th)
    den = ep + em
    a = (Vf / sz * ep + sz * If) / den
    b = (Vf / sz * em - sz * If) / den
    return a, b


def s_params(sig: dict, lines: dict, excited, f, dt, thr=1e-12, modes=None, normal_dx=None):
    """sig: {(port, m): (V, I)}, lines: {(port, m): (r, q, z0)} of the
    certified channels; the others need modes {(port, m): (omega_c, eps_r)}
    and normal_dx {port: dx}."""
    ab = {k: waves(V, I, f, dt, lines.get(k), None if modes is None else modes.get(k),
                   None if normal_dx is None else normal_dx.get(k[0]))
          for k, (V, I) in sig.items()}
    a_exc = ab[excited][0]
    valid = np.abs(a_exc) >= thr * np.max(np.abs(a_exc))
    safe = np.where(valid, a_exc, 1.0)
    return {k: np.where(valid, b / safe, np.nan) for k, (_, b) in ab.items()}
