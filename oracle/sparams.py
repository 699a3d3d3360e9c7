"""Oracle: port-signal DFT and power-wave S-parameters (CPU, float64).

Restates src/magnelio/signals/signal_1d.py:60-118 (direct DFT
sum_n x_n e^{-2 pi j f n dt}, no dt factor) and
src/magnelio/post/modal_sparameters.py:118-242, :335-557: I is shifted by
the temporal half step e^{+j w dt/2}; certified DTBC channels use the exact
discrete wave impedance (TE: z0 s/rad, TM: z0 rad/s) and the discrete
half-cell factor lambda^{1/2} for the two-plane de-stagger (a massless TEM
chain has no static z0 and takes the line impedance); channels on Mur the
continuum wave impedance (TE j w mu0 / gamma, TM gamma / (j w eps), TEM the
line impedance) and the factor e^{gamma dx_n / 2} (mode.py:170-231);
S = b / a_excited with |a| < 1e-12 peak -> NaN.

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
    """(a, b) at the port plane of a TE, TM or TEM channel: certified (line =
    (r, q, z0); z0 None on a massless chain, whose Z is the channel's own) or
    on Mur (line None, normal_dx the port cell); mode = (omega_c, eps_r[,
    kind[, z_line]])."""
    from .ports import z_mode  # noqa: PLC0415

    kind = mode[2] if mode is not None and len(mode) > 2 else "TE"
    z_line = mode[3] if mode is not None and len(mode) > 3 else None
    w = 2 * math.pi * np.asarray(f, float)
    Vf = dft(V, f, dt)
    If = dft(I, f, dt) * np.exp(1j * w * dt / 2)
    if line is not None and line[2] is not None:
        Z = dtbc_wave_impedance(w * dt, line[1], line[2], kind)
    else:
        Z = np.array([z_mode(kind, float(om), mode[0], mode[1], z_line)[0] for om in w], complex)
    if line is not None:
        th = destagger_theta(w * dt, line[0], line[1])
    else:
        th = 0.5 * float(normal_dx) * np.array([z_mode(kind, float(om), mode[0], mode[1])[1] for om in w],
                                               complex)
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
