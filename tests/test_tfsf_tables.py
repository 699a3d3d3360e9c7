"""CPU checks of the TF/SF correction tables (host integer/float work).

The product derives the twelve E and twelve H face patches from one curl
rule and ships them as a per-target correction list plus a waveform table;
applied on the CPU exactly as the device kernel does (target by target,
terms in order, field-dtype arithmetic), one injection must equal the
oracle's literal face-by-face restatement of plane_wave.py bit for bit, for
every propagation axis, both precisions, a non-uniform grid and several
steps.
"""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from magnelio_amd import layout
from magnelio_amd import mesh as mm
from magnelio_amd.sources import PlaneWaveSource
from oracle import mesh as om
from oracle import tfsf as ot
from oracle import yee

CASES = [
    ((0, 0, 1), (1, 0, 0)),
    ((0, 0, -1), (0, 1, 0)),
    ((1, 0, 0), (0, 1, 0)),
    ((-1, 0, 0), (0, 0, 1)),
    ((0, 1, 0), (0, 0, 1)),
    ((0, -1, 0), (1, 0, 0)),
    ((0, 0, 1), (1, 1, 0)),  # two polarisation components: every face patch active
]


def _grid(seed=5, dims=(11, 9, 13)):
    rng = np.random.default_rng(seed)
    return [np.cumsum(np.r_[-2e-3, rng.uniform(0.8e-3, 1.2e-3, n)]) for n in dims]


def _emulate(table, vec_dev, n):
    """Host emulation of k_tfsf on one device-layout vector."""
    tg = table["targets"].numpy()
    ptr = table["ptr"].numpy()
    coef = table["coef"].numpy()
    did = table["did"].numpy()
    wave = table["wave"].numpy()
    v = vec_dev[tg].copy()
    cnt = np.diff(ptr)
    for r in range(int(cnt.max())):
        m = cnt > r
        q = ptr[:-1][m] + r
        v[m] = v[m] + (coef[q] * wave[n, did[q]])
    out = vec_dev.copy()
    out[tg] = v
    return out


@pytest.mark.parametrize("reach_end", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("direction,polarization", CASES)
def test_tables_match_oracle_injection(direction, polarization, dtype, reach_end):
    """reach_end: the box touches the max face of the domain on every axis
    (no H layer beyond it; the E retardation falls back to the node)."""
    x, y, z = _grid()
    dims = (x.size - 1, y.size - 1, z.size - 1)
    rng = np.random.default_rng(1)
    be = rng.uniform(0.0, 2e-2, yee.n_e(*dims))
    bh = rng.uniform(0.0, 2e-2, yee.n_h(*dims))
    be[rng.random(be.size) < 0.05] = 0.0  # PEC-like zeros
    dt = om.courant_dt(x, y, z, 0.95)
    corners = ((x[3], y[2], z[4]), (x[8], y[7], z[10]))
    if reach_end:
        corners = (corners[0], (x[-1], y[-1], z[-1]))
    src = PlaneWaveSource(direction=direction, polarization=polarization, corners=corners,
                          amplitude=1.7, f_max=60e9)
    grid = mm.GridLines(x, y, z)
    box = ot.snap_box(x, y, z, corners)
    assert src.snap_box(grid) == box
    orc = ot.PlaneWave(x, y, z, be, bh, dt, direction=direction, polarization=polarization, box=box,
                       amplitude=1.7, f_max=60e9, dtype=dtype)
    n_steps = 120
    solver = SimpleNamespace(
        mesh=SimpleNamespace(grid=grid), dims=dims, _np_dtype=dtype,
        e=torch.zeros((1, yee.n_e(*dims))), _coeffs=SimpleNamespace(beta_e=be, beta_h=bh),
        dt=dt, total_time_steps=n_steps, horizon=n_steps, nsets=1)
    src.attach(solver)
    # fields of the size of the corrections, so every addition shows
    e = (rng.standard_normal(yee.n_e(*dims)) * 1e-8).astype(dtype)
    h = (rng.standard_normal(yee.n_h(*dims)) * 1e-8).astype(dtype)
    for n in (0, 41, 63, 119):
        t = n * dt
        oe, oh = e.copy(), h.copy()
        orc.inject_e(oe, t + dt)
        orc.inject_h(oh, t + dt + dt / 2)
        ge = _emulate(src._tables[0], layout.to_device(e, *dims, "E"), n)
        gh = _emulate(src._tables[1], layout.to_device(h, *dims, "H"), n)
        assert np.array_equal(layout.to_reference(ge, *dims, "E"), oe), n
        assert np.array_equal(layout.to_reference(gh, *dims, "H"), oh), n
        if n in (41, 63):  # pulse inside the box: the step injected something
            assert not np.array_equal(oe, e) and not np.array_equal(oh, h)


def test_snap_box_matches_oracle():
    x, y, z = _grid()
    g = mm.GridLines(x, y, z)
    for corners in (None, ((x[2], None, z[3]), (x[9], y[5], np.inf)), ((x[7], y[6], z[9]), (x[1], y[1], z[2]))):
        assert PlaneWaveSource(corners=corners).snap_box(g) == ot.snap_box(x, y, z, corners)


def test_oblique_incidence_refused():
    x, y, z = _grid()
    src = PlaneWaveSource(direction=(1, 1, 0), polarization=(0, 0, 1))
    with pytest.raises(NotImplementedError):
        src.patches(mm.GridLines(x, y, z), np.zeros(yee.n_e(10, 8, 12)), np.zeros(yee.n_h(10, 8, 12)),
                    np.float64)
