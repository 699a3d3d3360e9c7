"""Oracle: Yee field layout and the fused curl/material updates (CPU, NumPy).

Layout follows the reference ``FieldState``
(src/_fields/field_arrays.py:190-196): flat ``e = [Ex|Ey|Ez]``,
``h = [Hx|Hy|Hz]``, every component in C order with z fastest::

    Ex (Nx, Ny+1, Nz+1)   Ey (Nx+1, Ny, Nz+1)   Ez (Nx+1, Ny+1, Nz)
    Hx (Nx+1, Ny, Nz)     Hy (Nx, Ny+1, Nz)     Hz (Nx, Ny, Nz+1)

Two operation orders are provided, both from
src/_operators/numba_kernels.py:

* ``"stencil"`` — the array-stencil tier (update_E_stencil :137-171,
  update_H_stencil :174-205): curl accumulated in the field dtype as
  ``(((0 + t1) - t2) - t3) + t4``, then ``E = (a*E) + (b*curl)``,
  ``H = (a*H) - (b*curl)``.  This is the tier the reference runs on a CPU
  without Numba.
* ``"numba"`` — the Numba fused tier (update_E_fused :36-93,
  update_H_fused :95-127): the E-side curl starts from the float64 literal
  ``0.0`` and accumulates in float64; ``a*E`` is formed in the field dtype,
  ``b*curl`` and the sum in float64, rounded once on store.  The H side is
  field-dtype arithmetic in the same order as the stencil tier.

Here the one-sided terms at the array bounds are produced by zero padding:
adding or subtracting an exact zero leaves a floating-point value unchanged,
so the padded form rounds exactly like the reference's guarded form.

Parity pins (tests/test_oracle_yee.py): brute-force per-edge loops on small
random grids (bit-exact), the discrete identity div(C^T h) = 0 on interior
nodes, energy conservation of the lossless leapfrog, and the exact +0.0 of
frozen (alpha = beta = 0) edges.
"""

from __future__ import annotations

import numpy as np


def shapes(nx: int, ny: int, nz: int) -> dict[str, tuple[int, int, int]]:
    """Component shapes of an nx x ny x nz cell grid (field_arrays.py:190-196)."""
    return {
        "Ex": (nx, ny + 1, nz + 1),
        "Ey": (nx + 1, ny, nz + 1),
        "Ez": (nx + 1, ny + 1, nz),
        "Hx": (nx + 1, ny, nz),
        "Hy": (nx, ny + 1, nz),
        "Hz": (nx, ny, nz + 1),
    }


def split(vec: np.ndarray, nx: int, ny: int, nz: int, side: str) -> tuple[np.ndarray, ...]:
    """Zero-copy component views of a flat E (side='E') or H (side='H') vector."""
    sh = shapes(nx, ny, nz)
    names = ("Ex", "Ey", "Ez") if side == "E" else ("Hx", "Hy", "Hz")
    out = []
    off = 0
    for n in names:
        size = int(np.prod(sh[n]))
        out.append(vec[off : off + size].reshape(sh[n]))
        off += size
    if off != vec.size:
        raise ValueError(f"vector of length {vec.size} does not match grid {nx}x{ny}x{nz}")
    return tuple(out)


def n_e(nx: int, ny: int, nz: int) -> int:
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


def n_h(nx: int, ny: int, nz: int) -> int:
    return (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)


def _pad(a: np.ndarray, axis: int) -> np.ndarray:
    """Zero plane on both ends of ``axis`` (for one-sided boundary terms)."""
    width = [(0, 0)] * a.ndim
    width[axis] = (1, 1)
    return np.pad(a, width)


def _slab(a: np.ndarray, axis: int, lo: int, hi: int) -> np.ndarray:
    idx = [slice(None)] * a.ndim
    idx[axis] = slice(lo, hi)
    return a[tuple(idx)]


def _diff_dual(hp: np.ndarray, axis: int, n: int) -> tuple[np.ndarray, np.ndarray]:
    """(forward, backward) terms of a padded H component along ``axis``.

    ``hp`` is H zero-padded on ``axis``; for node index m = 0..n the forward
    term is H[m] (zero at m = n) and the backward term H[m-1] (zero at m = 0).
    """
    return _slab(hp, axis, 1, n + 2), _slab(hp, axis, 0, n + 1)


def curl_dual(e_shape_nodes, Hx, Hy, Hz, dtype, accumulate64: bool):
    """(C^T h) for the three E components; returns (cEx, cEy, cEz).

    Term order per component (numba_kernels.py:45-91 / :146-170):
      Ex: +Hz[j] -Hz[j-1] -Hy[k] +Hy[k-1]
      Ey: +Hx[k] -Hx[k-1] -Hz[i] +Hz[i-1]
      Ez: +Hy[i] -Hy[i-1] -Hx[j] +Hx[j-1]
    """
    nx, ny, nz = e_shape_nodes
    acc = np.float64 if accumulate64 else dtype

    def a(v):
        return v.astype(acc, copy=False)

    # Ex (nx, ny+1, nz+1)
    hz_f, hz_b = _diff_dual(_pad(a(Hz), 1), 1, ny)  # j terms, shape (nx, ny+1, nz+1)
    hy_f, hy_b = _diff_dual(_pad(a(Hy), 2), 2, nz)  # k terms, shape (nx, ny+1, nz+1)
    zero = np.zeros((), dtype=acc)
    cEx = (((zero + hz_f) - hz_b) - hy_f) + hy_b
    # Ey (nx+1, ny, nz+1)
    hx_f, hx_b = _diff_dual(_pad(a(Hx), 2), 2, nz)
    hz_fi, hz_bi = _diff_dual(_pad(a(Hz), 0), 0, nx)
    cEy = (((zero + hx_f) - hx_b) - hz_fi) + hz_bi
    # Ez (nx+1, ny+1, nz)
    hy_fi, hy_bi = _diff_dual(_pad(a(Hy), 0), 0, nx)
    hx_fj, hx_bj = _diff_dual(_pad(a(Hx), 1), 1, ny)
    cEz = (((zero + hy_fi) - hy_bi) - hx_fj) + hx_bj
    return cEx, cEy, cEz


def curl_primal(Ex, Ey, Ez):
    """(C e) for the three H components (numba_kernels.py:104-125), field dtype."""
    cHx = ((Ez[:, 1:, :] - Ez[:, :-1, :]) - Ey[:, :, 1:]) + Ey[:, :, :-1]
    cHy = ((Ex[:, :, 1:] - Ex[:, :, :-1]) - Ez[1:, :, :]) + Ez[:-1, :, :]
    cHz = ((Ey[1:, :, :] - Ey[:-1, :, :]) - Ex[:, 1:, :]) + Ex[:, :-1, :]
    return cHx, cHy, cHz


def update_e(e, h, alpha_e, beta_e, nx, ny, nz, order: str = "stencil") -> None:
    """E <- aE*E + bE*(C^T h), in place on the flat vector ``e``."""
    dtype = e.dtype
    Ex, Ey, Ez = split(e, nx, ny, nz, "E")
    Hx, Hy, Hz = split(h, nx, ny, nz, "H")
    aEx, aEy, aEz = split(alpha_e, nx, ny, nz, "E")
    bEx, bEy, bEz = split(beta_e, nx, ny, nz, "E")
    acc64 = order == "numba" and dtype == np.float32
    if order not in ("stencil", "numba"):
        raise ValueError(f"order must be 'stencil' or 'numba', got {order!r}")
    curls = curl_dual((nx, ny, nz), Hx, Hy, Hz, dtype, acc64)
    for F, a_, b_, c in zip((Ex, Ey, Ez), (aEx, aEy, aEz), (bEx, bEy, bEz), curls):
        if acc64:
            F[...] = ((a_ * F).astype(np.float64) + b_.astype(np.float64) * c).astype(dtype)
        else:
            F[...] = (a_ * F) + (b_ * c)


def update_h(e, h, alpha_h, beta_h, nx, ny, nz) -> None:
    """H <- aH*H - bH*(C e), in place on the flat vector ``h``."""
    Ex, Ey, Ez = split(e, nx, ny, nz, "E")
    Hx, Hy, Hz = split(h, nx, ny, nz, "H")
    aHx, aHy, aHz = split(alpha_h, nx, ny, nz, "H")
    bHx, bHy, bHz = split(beta_h, nx, ny, nz, "H")
    for F, a_, b_, c in zip((Hx, Hy, Hz), (aHx, aHy, aHz), (bHx, bHy, bHz), curl_primal(Ex, Ey, Ez)):
        F[...] = (a_ * F) - (b_ * c)


def energy(e, h, m_eps, m_mu) -> float:
    """0.5*(sum(M_eps e e) + sum(M_mu h h)) with float64 accumulation
    (src/solver/fit_td.py:1014-1019)."""
    pe = (m_eps * e) * e
    ph = (m_mu * h) * h
    return 0.5 * (float(np.sum(pe, dtype=np.float64)) + float(np.sum(ph, dtype=np.float64)))
