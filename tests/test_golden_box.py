"""CPML box with a TF/SF plane wave and DFT monitors: the oracle against the
reference's own run (tests/golden/box_tfsf_dft.npz, written by
tests/golden/make_box_fixtures.py from the reference's NumPy tier).

Graded 18 x 16 x 20 grid, two dielectric bricks, 4-layer CPML on all six
faces, +z x-polarised Gaussian TF/SF, a region E/H DFT monitor and a
far-field Huygens box 2 cells inside the absorber, 5 frequencies, 150 steps.
The oracle (oracle.cpml, oracle.tfsf, oracle.monitors) reproduces the
reference's fields and every accumulator's bins: float64 within 1e-12,
float32 bit for bit in the fields.
"""

import numpy as np
import pytest

import cases


@pytest.mark.parametrize("dtype,prec", [(np.float64, "double"), (np.float32, "single")])
def test_oracle_box_matches_reference(dtype, prec):
    f, mk = cases.golden_box()
    orc, region, far = cases.golden_box_oracle(dtype)
    orc.run(mk.STEPS)
    p = prec + "_"
    for comp, got in (("e", orc.e), ("h", orc.h)):
        want = f[p + comp]
        if dtype == np.float32:
            assert np.array_equal(got.astype(np.float64), want), comp
        else:
            err = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert err <= 1e-12, (comp, err)
    gate = 1e-12 if dtype == np.float64 else 1e-6
    for c, b in region.bins.items():
        want = f[f"{p}region_{c}"]
        assert b.shape == want.shape
        assert np.abs(b - want).max() <= gate * np.abs(want).max(), c
    peak = max(np.abs(f[k]).max() for k in f if k.startswith(p + "far_"))
    for face, comps in far.bins.items():
        for c, b in comps.items():
            want = f[f"{p}far_{face}_{c}"]
            assert b.shape == want.shape, (face, c)
            assert np.abs(b - want).max() <= gate * peak, (face, c)
