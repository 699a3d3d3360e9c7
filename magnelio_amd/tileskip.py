"""Dead-unit skipping for the curl kernels.

Reference: the static dead-tile analysis ``build_tile_skip_plan``
(src/magnelio/solver/_tile_skip.py:187-256) and its listed launches
(src/magnelio/_operators/numba_kernels.py:249-262, :459-474).  The rules
are the reference's:

* an E edge is dead iff alpha_E == 0 and beta_E == 0 (the final PEC mask,
  port flattening included): the update writes zero there forever;
* an H face is skippable iff alpha_H == 1 and beta_H == 0 (a no-op), or it
  is curl-dead (its four bounding E edges are dead, so its curl is zero
  forever) — here only where alpha_H == 1 as well, so the dense update
  h = alpha h - beta 0 leaves every value, not only a zero one, in place.
  The reference skips lossy curl-dead faces too and zeroes them on resume,
  which assumes zero initial fields; the stricter rule keeps skipping equal
  to the dense launch for any field state;
* configurations with TF/SF sources run dense (their H injection is scaled
  by beta_H, which is not zero inside metal).

The unit of skipping is this build's launch unit rather than the
reference's (2, 4, 32) tile: one work-group handles 256 consecutive nodes
(i, j) of the x-fastest node plane over ``kchunk`` z planes, every
component at those nodes.  A unit is launched when any element it updates
is live; the list holds unit ids ``kc * ntp + tp`` (ascending), the order
the dense launch walks them.  E edges written by a lumped element keep
their unit live (the element adds to them after the curl update, so the
curl update's zero write must happen first, as in the dense launch).
Dead E edges are zeroed once after a checkpoint load: the dense launch
writes zero there at the next E update, before anything reads them.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

WG = 256


@dataclass
class TilePlan:
    kchunk: int
    ntp: int
    nkc: int
    live_e: torch.Tensor  # int32 unit ids on the device
    live_h: torch.Tensor
    zero_e: torch.Tensor  # bool mask (device layout) of the dead E edges

    @property
    def units(self) -> int:
        return self.ntp * self.nkc

    @property
    def skipped_e(self) -> float:
        return 1.0 - self.live_e.numel() / self.units

    @property
    def skipped_h(self) -> float:
        return 1.0 - self.live_h.numel() / self.units


def _views_e(v, nx, ny, nz):
    """Device-layout E vector -> (Ex, Ey, Ez) views indexed [k, j, i]."""
    n0, n1 = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    return (v[:n0].view(nz + 1, ny + 1, nx), v[n0:n0 + n1].view(nz + 1, ny, nx + 1),
            v[n0 + n1:].view(nz, ny + 1, nx + 1))


def _views_h(v, nx, ny, nz):
    n0, n1 = (nx + 1) * ny * nz, nx * (ny + 1) * nz
    return (v[:n0].view(nz, ny, nx + 1), v[n0:n0 + n1].view(nz, ny + 1, nx),
            v[n0 + n1:].view(nz + 1, ny, nx))


def _node_grid(parts, nx, ny, nz):
    """OR of per-component masks placed on the (nz+1, ny+1, nx+1) node grid."""
    dev = parts[0].device
    g = torch.zeros((nz + 1, ny + 1, nx + 1), dtype=torch.bool, device=dev)
    for p in parts:
        kk, jj, ii = p.shape
        g[:kk, :jj, :ii] |= p
    return g


def _units(node, kchunk):
    """Unit liveness (nkc * ntp,) of a node-grid mask, unit id kc * ntp + tp."""
    nz1, ny1, nx1 = node.shape
    ntp = (nx1 * ny1 + WG - 1) // WG
    nkc = (nz1 + kchunk - 1) // kchunk
    flat = torch.zeros((nkc * kchunk, ntp * WG), dtype=torch.bool, device=node.device)
    flat[:nz1, :nx1 * ny1] = node.reshape(nz1, nx1 * ny1)
    u = flat.view(nkc, kchunk, ntp, WG).any(dim=3).any(dim=1)
    return u.reshape(-1), ntp, nkc


def build_tile_plan(alpha_e, beta_e, alpha_h, beta_h, dims, kchunk: int,
                    force_live_e=None) -> TilePlan:
    """Live-unit lists of the E and H launches from the device-layout
    coefficients (any torch device); ``force_live_e``: device-layout E
    indices whose units must run (lumped-element edges)."""
    nx, ny, nz = dims
    dead_e = (alpha_e == 0) & (beta_e == 0)
    ex, ey, ez = _views_e(dead_e, nx, ny, nz)
    # curl-dead H faces (_tile_skip.py:_curl_dead, on [k, j, i] views)
    cd = (ey[:-1] & ey[1:] & ez[:, :-1, :] & ez[:, 1:, :],       # Hx (nz, ny, nx+1)
          ez[:, :, :-1] & ez[:, :, 1:] & ex[:-1] & ex[1:],        # Hy (nz, ny+1, nx)
          ex[:, :-1, :] & ex[:, 1:, :] & ey[:, :, :-1] & ey[:, :, 1:])  # Hz (nz+1, ny, nx)
    noop = (alpha_h == 1) & (beta_h == 0)
    nh = _views_h(noop, nx, ny, nz)
    live_e_el = ~dead_e
    if force_live_e is not None and len(force_live_e):
        idx = torch.as_tensor(force_live_e, dtype=torch.int64, device=alpha_e.device)
        live_e_el = live_e_el.clone()
        live_e_el[idx] = True
    ue, ntp, nkc = _units(_node_grid(_views_e(live_e_el, nx, ny, nz), nx, ny, nz), kchunk)
    unit_h = _views_h(alpha_h == 1, nx, ny, nz)
    live_h_parts = [~(n | (c & u)) for n, c, u in zip(nh, cd, unit_h)]
    uh, _, _ = _units(_node_grid(live_h_parts, nx, ny, nz), kchunk)
    return TilePlan(kchunk, ntp, nkc, torch.nonzero(ue).reshape(-1).to(torch.int32),
                    torch.nonzero(uh).reshape(-1).to(torch.int32), dead_e)
