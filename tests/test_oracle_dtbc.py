"""Pins for the oracle's DTBC (CPU only).

Mirrors the reference's own gates (tests/unit/test_dtbc.py): the kernel
reconstructs the symbol on the contour, its leading Laurent coefficients
follow from lambda ~ r^2/z, the advance equals a brute-force convolution,
a terminated chain is transparent against a long unterminated chain, and
the incident-source path (u_inc, ghost) is part of the brute-force
convolution check.
"""

import math

import numpy as np
import pytest

from oracle import dtbc as od


@pytest.mark.parametrize("r,q", [(0.70, 0.0), (0.99, 0.0), (0.90, 0.20)])
def test_kernel_reconstructs_symbol(r, q):
    """The reference's gate (test_dtbc.py:97-104): length-4096 kernel, |z|=1.05."""
    ker = od.dtbc_kernel(r, q, 4096)
    z = 1.05 * np.exp(1j * np.linspace(0.0, np.pi, 17))
    m = np.arange(ker.size)
    approx = np.array([np.sum(ker * zz ** (-m)) for zz in z])
    assert np.max(np.abs(approx - od.lambda_symbol(z, r, q))) < 1e-12


def test_kernel_leading_terms():
    r, q = 0.8, 0.2
    ker = od.dtbc_kernel(r, q, 1024)
    assert abs(ker[0]) < 1e-12
    # lambda(z) = r^2/z + O(z^-2) as z -> infinity
    assert ker[1] == pytest.approx(r * r, rel=1e-10)


def test_symbol_is_outgoing_root():
    z = 1.3 * np.exp(1j * np.linspace(0, 2 * np.pi, 11))
    r, q = 0.6, 0.4
    lam = od.lambda_symbol(z, r, q)
    a = 1.0 + (z - 2.0 + 1.0 / z + q * q) / (2.0 * r * r)
    assert np.max(np.abs(lam * lam - 2 * a * lam + 1)) < 1e-12
    assert np.all(np.abs(lam) <= 1.0 + 1e-12)


def test_advance_matches_brute_force_convolution():
    rng = np.random.default_rng(0)
    r, q = 0.85, 0.1
    t = od.DTBC(r, q, n_kernel_init=64)
    ker = od.dtbc_kernel(r, q, 64)
    u_ints = rng.standard_normal(50)
    srcs = rng.standard_normal(50)
    u, u_prev, w, s = 0.0, 0.0, [], []
    for n in range(50):
        got = t.advance(u_ints[n], srcs[n])
        u_inc = sum(ker[m] * s[n - m] for m in range(1, n + 1))
        conv_w = sum(ker[m] * w[n - m] for m in range(1, n + 1))
        w.append(u - u_inc)
        s.append(srcs[n])
        u_new = (2 - 2 * r * r - q * q) * u - u_prev + r * r * (srcs[n] + conv_w + u_ints[n])
        u_prev, u = u, u_new
        assert got == pytest.approx(u_new, rel=1e-12, abs=1e-14)


def test_kernel_grows_by_doubling():
    t = od.DTBC(0.9, 0.0, n_kernel_init=16)
    for _ in range(40):
        t.advance(0.1, 0.0)
    assert t.nk == 64
    assert od.kernel_length_for_step(4095) == 4096
    assert od.kernel_length_for_step(4096) == 8192


def _chain(n_cells, steps, r, q, terminated):
    """Explicit leapfrog chain u_k (k=0..K) driven at k=0 by a pulse; the
    right end k=K is either DTBC-terminated or part of a long line."""
    K = n_cells
    u = np.zeros(K + 1)
    up = np.zeros(K + 1)
    term = od.DTBC(r, q) if terminated else None
    rec = []
    for n in range(steps):
        src = math.exp(-(((n - 40) / 10.0) ** 2))
        new = np.zeros_like(u)
        new[1:K] = (2 - 2 * r * r - q * q) * u[1:K] - up[1:K] + r * r * (u[2:] + u[:-2])[: K - 1]
        new[0] = src
        if terminated:
            new[K] = term.advance(u[K - 1], 0.0)
        else:
            new[K] = (2 - 2 * r * r - q * q) * u[K] - up[K] + r * r * u[K - 1]
        up, u = u, new
        rec.append(u[K // 2])
    return np.array(rec)


def test_terminated_chain_is_transparent():
    r, q = 0.9, 0.15
    steps = 400
    short = _chain(60, steps, r, q, terminated=True)
    long_ref = _chain(60 + 2 * steps, steps, r, q, terminated=False)
    # same probe position k = 30 in both chains (long chain probe index differs)
    K = 60 + 2 * steps
    u = np.zeros(K + 1)
    up = np.zeros(K + 1)
    rec = []
    for n in range(steps):
        src = math.exp(-(((n - 40) / 10.0) ** 2))
        new = np.zeros_like(u)
        new[1:K] = (2 - 2 * r * r - q * q) * u[1:K] - up[1:K] + r * r * (u[2:] + u[:-2])[: K - 1]
        new[0] = src
        up, u = u, new
        rec.append(u[30])
    ref = np.array(rec)
    err = np.max(np.abs(short - ref)) / np.max(np.abs(ref))
    assert err < 1e-5  # -100 dB
    del long_ref


def test_product_packed_kernels_follow_the_schedule():
    """The device's packed kernel table: version v (length init * 2^v) at
    offset init (2^v - 1), each equal to the oracle's kernel of that length,
    enough versions for the run; a first length below 2 is refused."""
    from magnelio_amd import dtbc as pd
    from magnelio_amd.solver import DevicePorts

    r, q, init, steps = 0.61, 0.03, 64, 300
    packed = pd.packed_kernels(r, q, steps, init)
    lengths = pd.kernel_versions(steps, init)
    assert lengths == [64, 128, 256, 512]
    off = 0
    for n in lengths:
        np.testing.assert_allclose(packed[off: off + n], od.dtbc_kernel(r, q, n), rtol=0, atol=1e-15)
        off += n
    assert packed.size == off
    with pytest.raises(ValueError, match="kernel_init"):
        DevicePorts([], (4, 4, 4), 1, 10, "cpu", [None], 1e-12, kernel_init=1)
