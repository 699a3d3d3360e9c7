"""Pins for the oracle's curl/material updates (CPU only).

* brute force: per-edge Python loops with the reference's guarded, ordered
  accumulation (src/magnelio/_operators/numba_kernels.py:36-127) — the
  oracle's padded vector form must agree bit for bit, in both operation
  orders and both precisions;
* discrete identity: the dual divergence of C^T h vanishes on interior nodes
  (C C^T structure), so a dropped or mis-signed term fails;
* energy: the lossless leapfrog conserves the discrete energy
  0.5 (e.M_eps e + h.M_mu h_avg) to round-off over many steps (test_fit_td
  :100-150 in the reference uses the same invariant);
* frozen edges (alpha = beta = 0) stay exactly +0.0.
"""

import numpy as np
import pytest

from oracle import yee


def _rand_state(rng, nx, ny, nz, dtype):
    ne, nh = yee.n_e(nx, ny, nz), yee.n_h(nx, ny, nz)
    e = rng.standard_normal(ne).astype(dtype)
    h = rng.standard_normal(nh).astype(dtype)
    ae = rng.uniform(0.5, 1.0, ne).astype(dtype)
    be = rng.uniform(0.1, 2.0, ne).astype(dtype)
    ah = rng.uniform(0.5, 1.0, nh).astype(dtype)
    bh = rng.uniform(0.1, 2.0, nh).astype(dtype)
    return e, h, ae, be, ah, bh


def _brute_e(e, h, ae, be, nx, ny, nz, order):
    dt = e.dtype
    Ex, Ey, Ez = (c.copy() for c in yee.split(e, nx, ny, nz, "E"))
    Hx, Hy, Hz = yee.split(h, nx, ny, nz, "H")
    aEx, aEy, aEz = yee.split(ae, nx, ny, nz, "E")
    bEx, bEy, bEz = yee.split(be, nx, ny, nz, "E")
    acc = np.float64 if (order == "numba") else dt.type

    def fin(a, f, b, c):
        if acc is np.float64 and dt == np.float32:
            return dt.type(np.float64(dt.type(a * f)) + np.float64(b) * c)
        return dt.type(dt.type(a * f) + dt.type(b * c))

    for i in range(nx):
        for j in range(ny + 1):
            for k in range(nz + 1):
                c = acc(0.0)
                if j < ny:
                    c = acc(c + acc(Hz[i, j, k]))
                if j > 0:
                    c = acc(c - acc(Hz[i, j - 1, k]))
                if k < nz:
                    c = acc(c - acc(Hy[i, j, k]))
                if k > 0:
                    c = acc(c + acc(Hy[i, j, k - 1]))
                Ex[i, j, k] = fin(aEx[i, j, k], Ex[i, j, k], bEx[i, j, k], c)
    for i in range(nx + 1):
        for j in range(ny):
            for k in range(nz + 1):
                c = acc(0.0)
                if k < nz:
                    c = acc(c + acc(Hx[i, j, k]))
                if k > 0:
                    c = acc(c - acc(Hx[i, j, k - 1]))
                if i < nx:
                    c = acc(c - acc(Hz[i, j, k]))
                if i > 0:
                    c = acc(c + acc(Hz[i - 1, j, k]))
                Ey[i, j, k] = fin(aEy[i, j, k], Ey[i, j, k], bEy[i, j, k], c)
    for i in range(nx + 1):
        for j in range(ny + 1):
            for k in range(nz):
                c = acc(0.0)
                if i < nx:
                    c = acc(c + acc(Hy[i, j, k]))
                if i > 0:
                    c = acc(c - acc(Hy[i - 1, j, k]))
                if j < ny:
                    c = acc(c - acc(Hx[i, j, k]))
                if j > 0:
                    c = acc(c + acc(Hx[i, j - 1, k]))
                Ez[i, j, k] = fin(aEz[i, j, k], Ez[i, j, k], bEz[i, j, k], c)
    return np.concatenate([Ex.ravel(), Ey.ravel(), Ez.ravel()])


def _brute_h(e, h, ah, bh, nx, ny, nz):
    t = h.dtype.type
    Ex, Ey, Ez = yee.split(e, nx, ny, nz, "E")
    Hx, Hy, Hz = (c.copy() for c in yee.split(h, nx, ny, nz, "H"))
    aHx, aHy, aHz = yee.split(ah, nx, ny, nz, "H")
    bHx, bHy, bHz = yee.split(bh, nx, ny, nz, "H")
    for i in range(nx + 1):
        for j in range(ny):
            for k in range(nz):
                c = t(t(t(Ez[i, j + 1, k] - Ez[i, j, k]) - Ey[i, j, k + 1]) + Ey[i, j, k])
                Hx[i, j, k] = t(t(aHx[i, j, k] * Hx[i, j, k]) - t(bHx[i, j, k] * c))
    for i in range(nx):
        for j in range(ny + 1):
            for k in range(nz):
                c = t(t(t(Ex[i, j, k + 1] - Ex[i, j, k]) - Ez[i + 1, j, k]) + Ez[i, j, k])
                Hy[i, j, k] = t(t(aHy[i, j, k] * Hy[i, j, k]) - t(bHy[i, j, k] * c))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz + 1):
                c = t(t(t(Ey[i + 1, j, k] - Ey[i, j, k]) - Ex[i, j + 1, k]) + Ex[i, j, k])
                Hz[i, j, k] = t(t(aHz[i, j, k] * Hz[i, j, k]) - t(bHz[i, j, k] * c))
    return np.concatenate([Hx.ravel(), Hy.ravel(), Hz.ravel()])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("order", ["stencil", "numba"])
@pytest.mark.parametrize("dims", [(3, 4, 5), (1, 1, 1), (2, 1, 3)])
def test_update_e_matches_brute_force(dtype, order, dims):
    rng = np.random.default_rng(7)
    nx, ny, nz = dims
    e, h, ae, be, _, _ = _rand_state(rng, nx, ny, nz, dtype)
    want = _brute_e(e, h, ae, be, nx, ny, nz, order)
    got = e.copy()
    yee.update_e(got, h, ae, be, nx, ny, nz, order=order)
    assert got.dtype == dtype
    assert np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("dims", [(3, 4, 5), (1, 1, 1), (2, 1, 3)])
def test_update_h_matches_brute_force(dtype, dims):
    rng = np.random.default_rng(11)
    nx, ny, nz = dims
    e, h, _, _, ah, bh = _rand_state(rng, nx, ny, nz, dtype)
    want = _brute_h(e, h, ah, bh, nx, ny, nz)
    got = h.copy()
    yee.update_h(e, got, ah, bh, nx, ny, nz)
    assert np.array_equal(got, want)


def test_numba_order_differs_from_stencil_in_float32():
    """The two orders are distinguishable, so each test above pins its own."""
    rng = np.random.default_rng(3)
    nx, ny, nz = 6, 5, 7
    e, h, ae, be, _, _ = _rand_state(rng, nx, ny, nz, np.float32)
    a = e.copy()
    b = e.copy()
    yee.update_e(a, h, ae, be, nx, ny, nz, order="stencil")
    yee.update_e(b, h, ae, be, nx, ny, nz, order="numba")
    assert not np.array_equal(a, b)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-5)


def test_dual_curl_is_divergence_free_on_interior_nodes():
    """div(C^T h) = 0 at interior nodes: with aE = 0, bE = 1 the E update
    returns C^T h exactly; its node divergence telescopes to zero."""
    rng = np.random.default_rng(5)
    nx, ny, nz = 5, 6, 4
    h = rng.standard_normal(yee.n_h(nx, ny, nz))
    e = np.zeros(yee.n_e(nx, ny, nz))
    yee.update_e(e, h, np.zeros_like(e), np.ones_like(e), nx, ny, nz)
    Ex, Ey, Ez = yee.split(e, nx, ny, nz, "E")
    div = (
        (Ex[1:, 1:-1, 1:-1] - Ex[:-1, 1:-1, 1:-1])
        + (Ey[1:-1, 1:, 1:-1] - Ey[1:-1, :-1, 1:-1])
        + (Ez[1:-1, 1:-1, 1:] - Ez[1:-1, 1:-1, :-1])
    )
    assert np.max(np.abs(div)) < 1e-12


def test_primal_curl_of_gradient_vanishes():
    """C grad(phi) = 0: edge voltages from a node potential have zero curl."""
    rng = np.random.default_rng(9)
    nx, ny, nz = 4, 5, 6
    phi = rng.standard_normal((nx + 1, ny + 1, nz + 1))
    Ex = phi[1:, :, :] - phi[:-1, :, :]
    Ey = phi[:, 1:, :] - phi[:, :-1, :]
    Ez = phi[:, :, 1:] - phi[:, :, :-1]
    e = np.concatenate([Ex.ravel(), Ey.ravel(), Ez.ravel()])
    h = np.zeros(yee.n_h(nx, ny, nz))
    yee.update_h(e, h, np.zeros_like(h), -np.ones_like(h), nx, ny, nz)
    assert np.max(np.abs(h)) < 1e-12


def test_lossless_leapfrog_conserves_discrete_energy():
    """Uniform vacuum box with PEC walls: the staggered energy invariant
    e.M_eps.e + h^{n+1/2}.M_mu.h^{n-1/2} stays constant to round-off."""
    from oracle import mesh as om

    nx, ny, nz = 6, 5, 7
    x = np.linspace(0, 6e-3, nx + 1)
    y = np.linspace(0, 5e-3, ny + 1)
    z = np.linspace(0, 7e-3, nz + 1)
    mid = np.zeros((nx, ny, nz), dtype=np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), nx, ny, nz, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    rng = np.random.default_rng(1)
    e = rng.standard_normal(me.size) * (~pec)
    h = rng.standard_normal(mm.size) * 1e-3
    vals = []
    for _ in range(200):
        h_old = h.copy()
        yee.update_e(e, h, ae, be, nx, ny, nz)
        yee.update_h(e, h, ah, bh, nx, ny, nz)
        vals.append(float(e @ (me * e) + h @ (mm * h_old)))
    vals = np.array(vals)
    assert np.ptp(vals) / abs(vals.mean()) < 1e-10


def test_frozen_edges_stay_positive_zero():
    rng = np.random.default_rng(2)
    nx, ny, nz = 4, 4, 4
    e, h, ae, be, _, _ = _rand_state(rng, nx, ny, nz, np.float32)
    frozen = rng.random(e.size) < 0.3
    e[frozen] = 0.0
    ae[frozen] = 0.0
    be[frozen] = 0.0
    yee.update_e(e, h, ae, be, nx, ny, nz)
    assert np.all(e[frozen] == 0.0)
    assert not np.any(np.signbit(e[frozen]))
