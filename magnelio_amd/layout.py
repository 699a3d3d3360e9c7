"""Reference (z-fastest) <-> device (x-fastest) layout of the Yee vectors.

The reference's flat vectors ``e = [Ex|Ey|Ez]`` and ``h = [Hx|Hy|Hz]`` keep
every component in C order with z fastest (src/_fields/field_arrays.py:42-196;
flat index formulas in src/mesh/indexing.py:17-63).  The device keeps the same
component blocks at the same offsets but orders each block x-fastest:
element (i, j, k) of a component of shape (n0, n1, n2) sits at
``k*n1*n0 + j*n0 + i`` inside its block.  Coalesced 64-lane accesses then run
along x, and a z-normal port plane is one contiguous run per component.

Every conversion here is an exact integer permutation (no arithmetic on the
values), so index maps such as PEC masks and port-plane gathers move between
the layouts bit-exactly.
"""

from __future__ import annotations

import numpy as np

E_NAMES = ("Ex", "Ey", "Ez")
H_NAMES = ("Hx", "Hy", "Hz")


def component_shapes(nx: int, ny: int, nz: int) -> dict[str, tuple[int, int, int]]:
    return {
        "Ex": (nx, ny + 1, nz + 1),
        "Ey": (nx + 1, ny, nz + 1),
        "Ez": (nx + 1, ny + 1, nz),
        "Hx": (nx + 1, ny, nz),
        "Hy": (nx, ny + 1, nz),
        "Hz": (nx, ny, nz + 1),
    }


def blocks(nx: int, ny: int, nz: int, side: str) -> list[tuple[str, int, tuple[int, int, int]]]:
    """[(name, offset, shape)] of the three component blocks of one side."""
    sh = component_shapes(nx, ny, nz)
    names = E_NAMES if side == "E" else H_NAMES
    out, off = [], 0
    for n in names:
        out.append((n, off, sh[n]))
        off += int(np.prod(sh[n]))
    return out


def vec_len(nx: int, ny: int, nz: int, side: str) -> int:
    return sum(int(np.prod(s)) for _, _, s in blocks(nx, ny, nz, side))


def to_device(vec: np.ndarray, nx: int, ny: int, nz: int, side: str) -> np.ndarray:
    """Reorder a reference-layout flat vector into the device layout."""
    vec = np.asarray(vec)
    if vec.shape[-1] != vec_len(nx, ny, nz, side):
        raise ValueError(f"vector length {vec.shape[-1]} does not match the {side} side")
    lead = vec.shape[:-1]
    out = np.empty_like(vec)
    for _, off, shp in blocks(nx, ny, nz, side):
        size = int(np.prod(shp))
        blk = vec[..., off : off + size].reshape(*lead, *shp)
        out[..., off : off + size] = np.moveaxis(blk, (-3, -2, -1), (-1, -2, -3)).reshape(*lead, size)
    return out


def to_reference(vec: np.ndarray, nx: int, ny: int, nz: int, side: str) -> np.ndarray:
    """Inverse of :func:`to_device`."""
    vec = np.asarray(vec)
    if vec.shape[-1] != vec_len(nx, ny, nz, side):
        raise ValueError(f"vector length {vec.shape[-1]} does not match the {side} side")
    lead = vec.shape[:-1]
    out = np.empty_like(vec)
    for _, off, (n0, n1, n2) in blocks(nx, ny, nz, side):
        size = n0 * n1 * n2
        blk = vec[..., off : off + size].reshape(*lead, n2, n1, n0)
        out[..., off : off + size] = np.moveaxis(blk, (-3, -2, -1), (-1, -2, -3)).reshape(*lead, size)
    return out


def index_to_device(idx: np.ndarray, nx: int, ny: int, nz: int, side: str) -> np.ndarray:
    """Map reference-layout flat indices to device-layout flat indices (int64)."""
    idx = np.asarray(idx, dtype=np.int64)
    out = np.empty_like(idx)
    total = vec_len(nx, ny, nz, side)
    if idx.size and (idx.min() < 0 or idx.max() >= total):
        raise IndexError(f"index out of range for the {side} vector of length {total}")
    for _, off, (n0, n1, n2) in blocks(nx, ny, nz, side):
        size = n0 * n1 * n2
        m = (idx >= off) & (idx < off + size)
        if not np.any(m):
            continue
        loc = idx[m] - off
        i, rem = np.divmod(loc, n1 * n2)
        j, k = np.divmod(rem, n2)
        out[m] = off + (k * n1 + j) * n0 + i
    return out


def index_to_reference(idx: np.ndarray, nx: int, ny: int, nz: int, side: str) -> np.ndarray:
    """Inverse of :func:`index_to_device`."""
    idx = np.asarray(idx, dtype=np.int64)
    out = np.empty_like(idx)
    for _, off, (n0, n1, n2) in blocks(nx, ny, nz, side):
        size = n0 * n1 * n2
        m = (idx >= off) & (idx < off + size)
        if not np.any(m):
            continue
        loc = idx[m] - off
        k, rem = np.divmod(loc, n1 * n0)
        j, i = np.divmod(rem, n0)
        out[m] = off + (i * n1 + j) * n2 + k
    return out
