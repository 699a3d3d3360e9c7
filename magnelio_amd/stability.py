"""Time-step limits: the geometric Courant step and the spectral CFL limit.

Reference: ``courant_dt`` (src/magnelio/solver/stability.py:44-93) and
``spectral_dt`` / ``_measure_lambda_max`` (:296-440):

    dt = safety * 2 / sqrt(lambda_max(D^-1/2 C^T M_mu^-1 C D^-1/2))

on the live E edges (not PEC, M_eps > 0), D = M_eps, 1/M_mu = 0 on frozen
faces; courant_dt when no edge is live.

The reference runs ARPACK on the host over a sparse curl matrix (its peak
memory on large meshes).  Here the operator never leaves the device: the
masses come from the device coefficient builder in float64 (uploaded from
the host construction for thin-wire meshes and hand-edited PEC masks), each
product is
one H and one E curl launch (alpha = 0, beta = -1/M_mu and 1) of the same
kernels that march the fields, and lambda_max comes from a Lanczos
recursion in float64 torch vectors (three of them; the tridiagonal matrix
is solved on the host every few steps) stopped when the residual bound of
the top Ritz value falls below ``tol`` relative.  Its value agrees with the
reference's eigsh (tol 1e-8) to that tolerance, not bit for bit.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from scipy.linalg import eigh_tridiagonal

from . import _lib, kernels
from . import mesh as _mesh
from .mesh import SAFETY_FACTORS, courant_dt

__all__ = ["courant_dt", "spectral_dt", "spectral_lambda_max"]


def _host_masses(mesh, dev):
    """The same from the host construction: meshes whose PEC mask or masses
    carry more than the cell materials (thin wires, hand-edited masks)."""
    from . import layout  # noqa: PLC0415

    dims = mesh.dims
    m_eps, m_mu = _mesh.build_m_eps(mesh), _mesh.build_m_mu(mesh)
    live = (~_mesh.pec_flat(mesh.pec_mask_edges, dims)) & (m_eps > 0)

    def up(a, side):
        return torch.from_numpy(np.ascontiguousarray(layout.to_device(a, *dims, side))).to(dev)

    return up(m_eps, "E"), up(m_mu, "H"), up(live, "E")


def _device_masses(mesh, dev):
    """float64 device-layout M_eps, M_mu and the live-E mask (beta_E != 0)."""
    if not getattr(mesh, "pec_rule", False) or getattr(mesh, "wire_corrections", None) is not None:
        return _host_masses(mesh, dev)
    dims = mesh.dims
    ne, nh = kernels.n_e(*dims), kernels.n_h(*dims)
    lib = mesh.material_library
    tab = np.zeros((max(lib) + 1, 16))
    for m, mat in lib.items():
        tab[m, 0:3] = mat.epsilon
        tab[m, 3:6] = mat.mu
        tab[m, 12] = 1.0 if mat.is_pec else 0.0
    g = mesh.grid
    geo = np.concatenate([g.dx, g.dy, g.dz, _mesh._avg_d(g.dx), _mesh._avg_d(g.dy), _mesh._avg_d(g.dz)])
    f64 = torch.float64
    out = [torch.empty(n, dtype=f64, device=dev) for n in (ne, ne, ne, nh, nh, nh)]
    mid = torch.from_numpy(np.ascontiguousarray(mesh.material_id, dtype=np.int32)).to(dev)
    t_tab, t_geo = torch.from_numpy(tab).to(dev), torch.from_numpy(geo).to(dev)
    walls = sum(1 << _mesh.FACES.index(f) for f in set(mesh.pec_faces))
    stream = torch.cuda.current_stream(dev)
    _lib.call("mgh_build_coefficients", 8, *dims, mid.data_ptr(), tab.shape[0], t_tab.data_ptr(),
              t_geo.data_ptr(), 1.0, float(_mesh.EPS0), float(_mesh.MU0), walls, 0,
              *(t.data_ptr() for t in out), int(stream.cuda_stream))
    stream.synchronize()
    _, beta_e, m_eps, _, _, m_mu = out
    return m_eps, m_mu, beta_e != 0


def spectral_lambda_max(mesh, device: str = "cuda:0", tol: float = 1e-10, max_iter: int = 5000,
                        check_every: int = 10) -> float | None:
    """lambda_max of the live operator (None when no E edge is live)."""
    _lib.load()
    dev = torch.device(device)
    dims = mesh.dims
    m_eps, m_mu, live = _device_masses(mesh, dev)
    if not bool(live.any()):
        return None
    inv_sqrt_d = torch.where(live, 1.0 / torch.sqrt(torch.where(live, m_eps, torch.ones_like(m_eps))),
                             torch.zeros_like(m_eps))
    mmu_inv = torch.where(m_mu > 0, 1.0 / torch.where(m_mu > 0, m_mu, torch.ones_like(m_mu)),
                          torch.zeros_like(m_mu))
    a_e, b_e = torch.zeros_like(m_eps), torch.ones_like(m_eps)
    a_h, b_h = torch.zeros_like(m_mu), -mmu_inv
    e = torch.zeros_like(m_eps)
    h = torch.zeros_like(m_mu)
    out = torch.zeros_like(m_eps)
    stream = torch.cuda.current_stream(dev)

    def apply(x):
        torch.mul(x, inv_sqrt_d, out=e)
        kernels.update_h(e, h, a_h, b_h, dims, stream=stream)    # h = C e / M_mu
        kernels.update_e(out, h, a_e, b_e, dims, stream=stream)  # out = 0 out + C^T h
        return out * inv_sqrt_d

    gen = torch.Generator(device=dev).manual_seed(0)
    v = torch.randn(m_eps.shape, generator=gen, dtype=torch.float64, device=dev) * live
    v /= torch.linalg.vector_norm(v)
    v_prev = torch.zeros_like(v)
    beta = 0.0
    alphas, betas = [], []
    theta = None
    for it in range(1, max_iter + 1):
        w = apply(v)
        alpha = float(torch.dot(w, v))
        w -= alpha * v
        if beta:
            w -= beta * v_prev
        beta = float(torch.linalg.vector_norm(w))
        alphas.append(alpha)
        betas.append(beta)
        if beta == 0.0 or it % check_every == 0 or it == max_iter:
            vals, vecs = eigh_tridiagonal(np.array(alphas), np.array(betas[:-1]),
                                          select="i", select_range=(it - 1, it - 1))
            theta = float(vals[0])
            if beta == 0.0 or abs(beta * vecs[-1, 0]) <= tol * abs(theta):
                return theta
        v_prev, v = v, w / beta
    return theta


def spectral_dt(mesh, accuracy: str = "normal", device: str = "cuda:0", tol: float = 1e-10) -> float:
    """Sharp CFL step safety * 2 / sqrt(lambda_max) (stability.py:296-356);
    the value is cached on the mesh like the reference's."""
    if accuracy not in SAFETY_FACTORS:
        raise ValueError(f"accuracy must be one of {list(SAFETY_FACTORS)!r}; got {accuracy!r}")
    safety = SAFETY_FACTORS[accuracy]
    lam = getattr(mesh, "_spectral_lambda_max", None)
    if lam is None:
        lam = spectral_lambda_max(mesh, device=device, tol=tol)
        if lam is None or not math.isfinite(lam) or lam <= 0.0:
            return courant_dt(mesh.grid, accuracy)
        mesh._spectral_lambda_max = lam
    return safety * 2.0 / math.sqrt(lam)
