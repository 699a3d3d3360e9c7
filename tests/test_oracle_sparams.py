"""Pins for the oracle's S-parameter chain (CPU only).

* DFT closed forms: a sampled sine and a constant (test_monitors.py:762-787
  in the reference uses the same pair);
* physics on a reduced WR-90 (24 x 12 x 48 cells, both ends DTBC-terminated,
  TE10 excited): the empty uniform guide is reflection-free and lossless,
  |S21| = 1 and |S11| at the numerical floor across the TE10 band, higher
  modes stay unexcited, and the S21 phase equals -beta_hat L from the
  discrete dispersion sin^2(w dt/2) = r^2 sin^2(beta_hat dz/2) + q^2/4
  (src/magnelio/ports/_modal/dtbc.py:58-66).
"""

import math

import numpy as np
import pytest

import cases
from oracle import sparams as osp


def test_dft_closed_forms():
    dt = 1e-12
    n = np.arange(2000)
    f0 = 37.5e9 / 3
    x = np.cos(2 * math.pi * f0 * n * dt)
    F = osp.dft(x, [f0], dt)[0]
    # sum cos^2-like: N/2 plus a small Dirichlet-kernel term
    assert abs(F.real - 1000) < 2.0 and abs(F.imag) < 2.0
    assert osp.dft(np.ones(100), [0.0], dt)[0] == pytest.approx(100.0)
    f1 = 1.0 / (50 * dt)
    assert abs(osp.dft(np.ones(100), [f1], dt)[0]) < 1e-10


@pytest.fixture(scope="module")
def small_c1():
    dims, size = (24, 12, 48), cases.C1_SIZE
    # a band clear of the TE10 cut-off: cut-off content rings algebraically
    # (zero group velocity) and would need a far longer record
    orc, ports, dt, _ = cases.c1_oracle(np.float64, dims=dims, size=size, band=(11e9, 20e9))
    orc.run(4000)
    return orc, ports, dt, dims, size


def test_reduced_wr90_is_matched_and_lossless(small_c1):
    orc, ports, dt, dims, size = small_c1
    p1, p2 = ports
    sig = {}
    lines = {}
    for p in ports:
        for m in range(p.n_modes):
            sig[(p.name, m)] = (np.array(p.V[m]), np.array(p.I[m]))
            lines[(p.name, m)] = p.line[m]
    f_c = p1.modes[0]["wc"] / (2 * math.pi)
    f = np.linspace(12e9, 19e9, 29)
    del f_c
    S = osp.s_params(sig, lines, ("port1", 0), f, dt)
    s21 = S[("port2", 0)]
    s11 = S[("port1", 0)]
    # limited by the cut-off ringing tail of the finite record (~1e-5)
    assert np.max(np.abs(np.abs(s21) - 1)) < 1e-4
    assert 20 * np.log10(np.max(np.abs(s11))) < -90
    for k in (("port2", 1), ("port2", 2), ("port1", 1), ("port1", 2)):
        assert 20 * np.log10(np.max(np.abs(S[k]))) < -90
    # discrete dispersion phase
    r, q, _ = p1.line[0]
    dz = size[2] / dims[2]
    w = 2 * math.pi * f
    sb = np.sqrt((np.sin(w * dt / 2) ** 2 - q * q / 4) / (r * r))
    beta = 2 * np.arcsin(sb) / dz
    L = size[2]
    ph_err = np.angle(s21 * np.exp(1j * beta * L))
    assert np.max(np.abs(ph_err)) < 1e-4
