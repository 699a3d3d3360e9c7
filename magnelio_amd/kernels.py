"""Python wrappers of the libmgh entry points over torch device tensors.

Device buffers are torch tensors on the ROCm device (PyTorch is only the
allocator and stream provider here); every call passes raw pointers, sizes
and the current HIP stream through the C-ABI.  Shapes and dtypes are
checked on the host before any launch.

Kernel-tier counterparts in the reference:
``update_E_fused_cuda`` / ``update_H_fused_cuda``
(src/_operators/numba_kernels.py:518-552): in place, float32/float64 only,
``TypeError`` for any other dtype (:483-488).
"""

from __future__ import annotations

import torch

from . import _lib

DEFAULT_KCHUNK = 2


def launch_shape(kchunk: int, sets: int = 0, band: int = 0) -> int:
    """The curl launches' shape word (the ``kchunk`` argument of every curl
    entry point): z planes marched per thread (bits 0-15), field sets per
    thread (bits 16-19; 0: the library's choice) and the XCD band in launch
    units (bits 20-27; 0: 8).  Results do not depend on it."""
    if not (1 <= kchunk <= 0xFFFF and 0 <= sets <= 3 and 0 <= band <= 0xFF):
        raise ValueError(f"bad launch shape kchunk={kchunk} sets={sets} band={band}")
    return int(kchunk) | (int(sets) << 16) | (int(band) << 20)


def shape_parts(word: int) -> tuple[int, int, int]:
    """(z planes per thread, sets per thread, band) of a shape word; 0 for a
    part left to the library."""
    return word & 0xFFFF, (word >> 16) & 0xF, (word >> 20) & 0xFF


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return 4
    if t.dtype == torch.float64:
        return 8
    raise TypeError(f"HIP fused kernels support float32/float64, got {t.dtype}")


def stream_handle(stream: torch.cuda.Stream | None = None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _check_dev(*ts: torch.Tensor) -> None:
    for t in ts:
        if not t.is_cuda:
            raise ValueError("HIP backend buffers must live on the GPU")
        if not t.is_contiguous():
            raise ValueError("HIP backend buffers must be contiguous")


def n_e(nx: int, ny: int, nz: int) -> int:
    return nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz


def n_h(nx: int, ny: int, nz: int) -> int:
    return (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)


def _cpml_desc(cpml, nsets: int, dtype) -> int | None:
    if cpml is None:
        return None
    if cpml.nsets != nsets:
        raise ValueError(f"CPML state holds {cpml.nsets} field sets, the launch {nsets}")
    if cpml.dtype != dtype:
        raise TypeError("CPML state and fields must share one dtype")
    return cpml.desc_ptr


def update_e(e: torch.Tensor, h: torch.Tensor, alpha_e: torch.Tensor, beta_e: torch.Tensor,
             dims: tuple[int, int, int], *, curl64: bool = False, kchunk: int = DEFAULT_KCHUNK,
             cpml=None, live: torch.Tensor | None = None, xcols: bool = True,
             stream: torch.cuda.Stream | None = None) -> None:
    """E phase on K field sets: ``e`` is (K, n_E) or (n_E,), ``h`` likewise.

    ``cpml`` is a :class:`magnelio_amd.boundaries.DeviceCPML` (psi slabs for
    the same K sets) or None: it rides in the curl launch, except the
    x-slab columns when the x slabs are not both at the domain ends, which
    follow in a launch of their own (``xcols=False`` leaves that one to
    :func:`cpml_xcols_e`); ``live``: int32 device list of the launch units to
    run (``tileskip.build_tile_plan``), None for all."""
    nx, ny, nz = dims
    ne, nh = n_e(nx, ny, nz), n_h(nx, ny, nz)
    e2 = e.view(-1, ne)
    h2 = h.view(-1, nh)
    k = e2.shape[0]
    if h2.shape[0] != k:
        raise ValueError(f"E has {k} field sets, H has {h2.shape[0]}")
    if alpha_e.numel() != ne or beta_e.numel() != ne:
        raise ValueError("alpha_E/beta_E length does not match the grid")
    code = _dtype_code(e)
    for t in (h, alpha_e, beta_e):
        if t.dtype != e.dtype:
            raise TypeError("fields and coefficients must share one dtype")
    _check_dev(e, h, alpha_e, beta_e)
    if live is not None:
        if cpml is not None:
            raise ValueError("listed launches take the CPML passes separately (cpml_e)")
        _check_live(live, e)
        _lib.call("mgh_update_e_tiles", code, nx, ny, nz, k, e.data_ptr(), ne, h.data_ptr(), nh,
                  alpha_e.data_ptr(), beta_e.data_ptr(), int(bool(curl64)), int(kchunk),
                  live.data_ptr(), live.numel(), stream_handle(stream))
        return
    desc = _cpml_desc(cpml, k, e.dtype)
    _lib.call("mgh_update_e_pml" if xcols else "mgh_update_e_pml_core", code, nx, ny, nz, k,
              e.data_ptr(), ne, h.data_ptr(), nh,
              alpha_e.data_ptr(), beta_e.data_ptr(), int(bool(curl64)), int(kchunk), desc,
              stream_handle(stream))


def cpml_e(e: torch.Tensor, h: torch.Tensor, beta_e: torch.Tensor, dims: tuple[int, int, int],
           cpml, *, stream: torch.cuda.Stream | None = None) -> None:
    """CPML slab passes of the E phase alone (after :func:`update_e` without ``cpml``)."""
    nx, ny, nz = dims
    ne, nh = n_e(nx, ny, nz), n_h(nx, ny, nz)
    k = e.view(-1, ne).shape[0]
    if h.view(-1, nh).shape[0] != k or beta_e.numel() != ne:
        raise ValueError("field sets or coefficient length do not match the grid")
    _check_dev(e, h, beta_e)
    _lib.call("mgh_cpml_e", _dtype_code(e), nx, ny, nz, k, e.data_ptr(), ne, h.data_ptr(), nh,
              beta_e.data_ptr(), _cpml_desc(cpml, k, e.dtype), stream_handle(stream))


def cpml_h(e: torch.Tensor, h: torch.Tensor, beta_h: torch.Tensor, dims: tuple[int, int, int],
           cpml, *, stream: torch.cuda.Stream | None = None) -> None:
    """CPML slab passes of the H phase alone (after :func:`update_h` without ``cpml``)."""
    nx, ny, nz = dims
    ne, nh = n_e(nx, ny, nz), n_h(nx, ny, nz)
    k = e.view(-1, ne).shape[0]
    if h.view(-1, nh).shape[0] != k or beta_h.numel() != nh:
        raise ValueError("field sets or coefficient length do not match the grid")
    _check_dev(e, h, beta_h)
    _lib.call("mgh_cpml_h", _dtype_code(h), nx, ny, nz, k, e.data_ptr(), ne, h.data_ptr(), nh,
              beta_h.data_ptr(), _cpml_desc(cpml, k, h.dtype), stream_handle(stream))


def cpml_xcols_e(e: torch.Tensor, h: torch.Tensor, beta_e: torch.Tensor,
                 dims: tuple[int, int, int], cpml, *,
                 stream: torch.cuda.Stream | None = None) -> None:
    """CPML of the x-slab columns, E phase (after :func:`update_e` with
    ``cpml`` and ``xcols=False``): every face, in order, on those elements."""
    _xcols("mgh_cpml_xcols_e", e, h, beta_e, dims, cpml, stream, e)


def cpml_xcols_h(e: torch.Tensor, h: torch.Tensor, beta_h: torch.Tensor,
                 dims: tuple[int, int, int], cpml, *,
                 stream: torch.cuda.Stream | None = None) -> None:
    """CPML of the x-slab columns, H phase (after :func:`update_h` with
    ``cpml`` and ``xcols=False``)."""
    _xcols("mgh_cpml_xcols_h", e, h, beta_h, dims, cpml, stream, h)


def _xcols(name, e, h, beta, dims, cpml, stream, side) -> None:
    nx, ny, nz = dims
    ne, nh = n_e(nx, ny, nz), n_h(nx, ny, nz)
    k = e.view(-1, ne).shape[0]
    if h.view(-1, nh).shape[0] != k or beta.numel() != side.view(k, -1).shape[1]:
        raise ValueError("field sets or coefficient length do not match the grid")
    _check_dev(e, h, beta)
    _lib.call(name, _dtype_code(side), nx, ny, nz, k, e.data_ptr(), ne, h.data_ptr(), nh,
              beta.data_ptr(), _cpml_desc(cpml, k, side.dtype), stream_handle(stream))


def update_h(e: torch.Tensor, h: torch.Tensor, alpha_h: torch.Tensor, beta_h: torch.Tensor,
             dims: tuple[int, int, int], *, kchunk: int = DEFAULT_KCHUNK, cpml=None,
             live: torch.Tensor | None = None, xcols: bool = True,
             stream: torch.cuda.Stream | None = None) -> None:
    """H phase on K field sets (same conventions as :func:`update_e`)."""
    nx, ny, nz = dims
    ne, nh = n_e(nx, ny, nz), n_h(nx, ny, nz)
    e2 = e.view(-1, ne)
    h2 = h.view(-1, nh)
    k = e2.shape[0]
    if h2.shape[0] != k:
        raise ValueError(f"E has {k} field sets, H has {h2.shape[0]}")
    if alpha_h.numel() != nh or beta_h.numel() != nh:
        raise ValueError("alpha_H/beta_H length does not match the grid")
    code = _dtype_code(h)
    for t in (e, alpha_h, beta_h):
        if t.dtype != h.dtype:
            raise TypeError("fields and coefficients must share one dtype")
    _check_dev(e, h, alpha_h, beta_h)
    if live is not None:
        if cpml is not None:
            raise ValueError("listed launches take the CPML passes separately (cpml_h)")
        _check_live(live, h)
        _lib.call("mgh_update_h_tiles", code, nx, ny, nz, k, e.data_ptr(), ne, h.data_ptr(), nh,
                  alpha_h.data_ptr(), beta_h.data_ptr(), int(kchunk), live.data_ptr(), live.numel(),
                  stream_handle(stream))
        return
    desc = _cpml_desc(cpml, k, h.dtype)
    _lib.call("mgh_update_h_pml" if xcols else "mgh_update_h_pml_core", code, nx, ny, nz, k,
              e.data_ptr(), ne, h.data_ptr(), nh,
              alpha_h.data_ptr(), beta_h.data_ptr(), int(kchunk), desc, stream_handle(stream))


def _check_live(live: torch.Tensor, like: torch.Tensor) -> None:
    if live.dtype != torch.int32 or live.dim() != 1 or not live.is_contiguous():
        raise TypeError("live must be a contiguous 1-D int32 tensor")
    if live.device != like.device:
        raise ValueError("live list and fields must be on the same device")


class SqSum:
    """Reusable float64 weighted-square-sum reduction (device result)."""

    def __init__(self, nsets: int, device: torch.device):
        lib = _lib.load()
        self.nsets = nsets
        self.work = torch.empty(int(lib.mgh_weighted_sq_sum_workspace(nsets)), dtype=torch.float64,
                                device=device)

    def __call__(self, x: torch.Tensor, w: torch.Tensor, out: torch.Tensor, *,
                 accumulate: bool = False, scale: float = 1.0,
                 stream: torch.cuda.Stream | None = None) -> None:
        n = w.numel()
        x2 = x.view(-1, n)
        k = x2.shape[0]
        if k > self.nsets:
            raise ValueError(f"{k} sets exceed the reduction's capacity {self.nsets}")
        if out.dtype != torch.float64 or out.numel() < k:
            raise ValueError("out must be a float64 device tensor with one slot per set")
        if x.dtype != w.dtype:
            raise TypeError("values and weights must share one dtype")
        _check_dev(x, w, out)
        _lib.call("mgh_weighted_sq_sum", _dtype_code(x), n, k, x.data_ptr(), n, w.data_ptr(),
                  out.data_ptr(), int(accumulate), float(scale), self.work.data_ptr(),
                  self.work.numel(), stream_handle(stream))
