"""Oracle: the spectral CFL limit (CPU).  Test infrastructure only.

Restates spectral_dt / _measure_lambda_max (src/magnelio/solver/stability.py
:296-440, paths relative to the reference repository root):

    dt = safety * 2 / sqrt(lambda_max(D^-1/2 C^T M_mu^-1 C D^-1/2))

on the live E edges (not PEC, M_eps > 0), D = M_eps, with 1/M_mu = 0 on
frozen faces; lambda_max from scipy's eigsh (which="LA", tol 1e-8, start
vector default_rng(0).standard_normal(n_live)), the Gershgorin row-sum bound
of the same operator when ARPACK fails or fewer than 8 edges are live, and
courant_dt when no edge is live.  The matvec is the oracle's own curl pair
(yee.update_h / yee.update_e with alpha = 0).

Pinned by tests/test_oracle_stability.py: on a uniform PEC box the largest
eigenvalue is the closed form
    (4 / (mu0 eps0)) sum_a sin^2((N_a - 1) pi / (2 N_a)) / d_a^2
of the discrete Maxwell operator, and the Gershgorin bound is never below it.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse.linalg as spla

from . import yee


def lambda_max(dims, m_eps, m_mu, pec_flat):
    """(lambda_max, gershgorin bound) of the live operator, or (None, None)."""
    nx, ny, nz = dims
    live = (~np.asarray(pec_flat, bool)) & (m_eps > 0)
    if not live.any():
        return None, None
    idx = np.nonzero(live)[0]
    inv_sqrt_d = 1.0 / np.sqrt(m_eps[idx])
    mmu_inv = np.where(m_mu > 0, 1.0 / np.where(m_mu > 0, m_mu, 1.0), 0.0)
    ne, nh = yee.n_e(nx, ny, nz), yee.n_h(nx, ny, nz)
    zero_e, one_e, zero_h = np.zeros(ne), np.ones(ne), np.zeros(nh)

    def apply(x, absolute=False):
        e = np.zeros(ne)
        e[idx] = np.asarray(x).ravel() * inv_sqrt_d
        h = np.zeros(nh)
        if absolute:
            # |C| products: the curl with every difference replaced by a sum
            h = _abs_curl_h(e, dims) * mmu_inv
            out = _abs_curl_e(h, dims)
        else:
            yee.update_h(e, h, zero_h, -mmu_inv, nx, ny, nz)
            out = np.zeros(ne)
            yee.update_e(out, h, zero_e, one_e, nx, ny, nz)
        return out[idx] * inv_sqrt_d

    upper = float(apply(np.ones(idx.size), absolute=True).max())
    if idx.size < 8:
        return upper, upper
    op = spla.LinearOperator((idx.size, idx.size), matvec=apply, dtype=np.float64)
    v0 = np.random.default_rng(0).standard_normal(idx.size)
    try:
        lam = float(spla.eigsh(op, k=1, which="LA", return_eigenvectors=False, tol=1e-8,
                               maxiter=10_000, v0=v0)[0])
    except spla.ArpackError:
        return upper, upper
    if not math.isfinite(lam) or lam <= 0.0:
        return upper, upper
    return lam, upper


def _abs_curl_h(e, dims):
    """|C| e: the primal curl with every term added (the row-sum bound)."""
    nx, ny, nz = dims
    Ex, Ey, Ez = (np.abs(a) for a in yee.split(e, nx, ny, nz, "E"))
    hx = Ez[:, 1:, :] + Ez[:, :-1, :] + Ey[:, :, 1:] + Ey[:, :, :-1]
    hy = Ex[:, :, 1:] + Ex[:, :, :-1] + Ez[1:, :, :] + Ez[:-1, :, :]
    hz = Ey[1:, :, :] + Ey[:-1, :, :] + Ex[:, 1:, :] + Ex[:, :-1, :]
    return np.concatenate([hx.ravel(), hy.ravel(), hz.ravel()])


def _abs_curl_e(h, dims):
    """|C|^T h: every face adjacent to an edge contributes with weight 1."""
    nx, ny, nz = dims
    Hx, Hy, Hz = yee.split(h, nx, ny, nz, "H")
    ex = np.zeros((nx, ny + 1, nz + 1))
    ey = np.zeros((nx + 1, ny, nz + 1))
    ez = np.zeros((nx + 1, ny + 1, nz))
    # Hz(i,j,k) touches Ex(i,j,k), Ex(i,j+1,k), Ey(i,j,k), Ey(i+1,j,k)
    ex[:, :-1, :] += Hz; ex[:, 1:, :] += Hz
    ey[:-1, :, :] += Hz; ey[1:, :, :] += Hz
    # Hy(i,j,k) touches Ex(i,j,k), Ex(i,j,k+1), Ez(i,j,k), Ez(i+1,j,k)
    ex[:, :, :-1] += Hy; ex[:, :, 1:] += Hy
    ez[:-1, :, :] += Hy; ez[1:, :, :] += Hy
    # Hx(i,j,k) touches Ey(i,j,k), Ey(i,j,k+1), Ez(i,j,k), Ez(i,j+1,k)
    ey[:, :, :-1] += Hx; ey[:, :, 1:] += Hx
    ez[:, :-1, :] += Hx; ez[:, 1:, :] += Hx
    return np.concatenate([ex.ravel(), ey.ravel(), ez.ravel()])


def box_lambda_max(dims, d, eps0, mu0):
    """Closed form of the largest eigenvalue on a uniform PEC box."""
    return (4.0 / (mu0 * eps0)) * sum(math.sin((n - 1) * math.pi / (2 * n)) ** 2 / dd ** 2
                                      for n, dd in zip(dims, d))
