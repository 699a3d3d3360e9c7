"""ctypes binding of ``libmgh.so`` (declared in include/mgh.h).

The library is loaded only on demand.  Before it is loaded, the HIP runtime
that PyTorch-ROCm ships is promoted to global symbol scope, because the
library is linked without a runtime of its own and must share torch's
(device pointers and streams cross the boundary as plain integers).

A missing or unloadable library raises :class:`HipBackendUnavailable`;
nothing in this package falls back to a CPU path.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmgh.so"

_c_int = ctypes.c_int
_c_ll = ctypes.c_longlong
_c_void_p = ctypes.c_void_p
_c_double = ctypes.c_double
_c_char_p = ctypes.c_char_p

# name -> (restype, argtypes); mirrors include/mgh.h
SIGNATURES: dict[str, tuple] = {
    "mgh_last_error": (_c_char_p, []),
    "mgh_version": (_c_int, []),
    "mgh_device_info": (
        _c_int,
        [_c_char_p, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_ll), _c_char_p, _c_int],
    ),
    "mgh_stream_probe": (_c_int, [_c_int, _c_int, _c_void_p, _c_void_p, _c_ll, _c_void_p]),
    "mgh_update_e": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_int, _c_void_p],
    ),
    "mgh_update_h": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_void_p],
    ),
    "mgh_update_e_pml": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_update_h_pml": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_update_e_pml_core": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_update_h_pml_core": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_cpml_xcols_e": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_void_p],
    ),
    "mgh_cpml_xcols_h": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_void_p],
    ),
    "mgh_cpml_e": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_void_p],
    ),
    "mgh_cpml_h": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_void_p],
    ),
    "mgh_port_step": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p, _c_int, _c_void_p,
         _c_void_p],
    ),
    "mgh_step_advance": (_c_int, [_c_void_p, _c_ll, _c_void_p]),
    "mgh_update_e_packed": (_c_int, [_c_int, _c_int, _c_void_p, _c_int, _c_void_p]),
    "mgh_update_h_packed": (_c_int, [_c_int, _c_int, _c_void_p, _c_int, _c_void_p]),
    "mgh_port_step_packed": (_c_int, [_c_int, _c_int, _c_void_p, _c_void_p, _c_void_p]),
    "mgh_step_advance_packed": (_c_int, [_c_int, _c_void_p, _c_ll, _c_void_p]),
    "mgh_weighted_sq_sum_workspace": (_c_ll, [_c_int]),
    "mgh_weighted_sq_sum": (
        _c_int,
        [_c_int, _c_ll, _c_int, _c_void_p, _c_ll, _c_void_p, _c_void_p, _c_int, _c_double,
         _c_void_p, _c_ll, _c_void_p],
    ),
    "mgh_weighted_sq_sum_noalloc": (
        _c_int,
        [_c_int, _c_ll, _c_int, _c_void_p, _c_ll, _c_void_p, _c_void_p, _c_int, _c_double,
         _c_void_p],
    ),
    "mgh_lumped_step": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_build_coefficients": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_double,
         _c_double, _c_double, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
         _c_void_p, _c_void_p, _c_void_p],
    ),
    "mgh_update_e_tiles": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_int, _c_void_p, _c_int, _c_void_p],
    ),
    "mgh_update_h_tiles": (
        _c_int,
        [_c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p,
         _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p],
    ),
    "mgh_ade_save": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_int, _c_void_p, _c_void_p, _c_void_p],
    ),
    "mgh_ade_update": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
         _c_void_p, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_sibc_update": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_int,
         _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int,
         _c_void_p, _c_void_p],
    ),
    "mgh_ntff": (
        _c_int,
        [_c_int, _c_void_p, _c_int, _c_void_p, _c_double, _c_double, _c_double, _c_double, _c_int,
         _c_void_p, _c_void_p, _c_void_p],
    ),
    "mgh_ntff_faces": (
        _c_int,
        [_c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int, _c_void_p, _c_int, _c_void_p,
         _c_double, _c_double, _c_double, _c_double, _c_void_p, _c_void_p, _c_void_p],
    ),
    "mgh_tfsf_apply": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_void_p,
         _c_void_p, _c_int, _c_ll, _c_void_p, _c_ll, _c_void_p],
    ),
    "mgh_dft_sample": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p, _c_void_p, _c_int,
         _c_void_p, _c_int, _c_int, _c_ll, _c_void_p, _c_ll, _c_void_p],
    ),
    "mgh_dft_sample_groups": (
        _c_int,
        [_c_int, _c_int, _c_void_p, _c_ll, _c_void_p, _c_ll, _c_void_p, _c_int, _c_int, _c_void_p,
         _c_int, _c_void_p, _c_int, _c_int, _c_ll, _c_void_p, _c_ll, _c_void_p],
    ),
    "mgh_dft_block": (
        _c_int,
        [_c_int, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p,
         _c_int, _c_int, _c_double, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_ll,
         _c_void_p],
    ),
    "mgh_dft_blocks": (
        _c_int,
        [_c_int, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p,
         _c_void_p, _c_int, _c_int, _c_double, _c_void_p, _c_int, _c_void_p, _c_void_p, _c_void_p,
         _c_ll, _c_void_p],
    ),
    "mgh_dft_blocks_range": (
        _c_int,
        [_c_int, _c_void_p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_ll,
         _c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_double, _c_void_p, _c_int, _c_void_p,
         _c_void_p, _c_int, _c_void_p],
    ),
    "mgh_dft_phase_table": (
        _c_int,
        [_c_void_p, _c_int, _c_double, _c_ll, _c_int, _c_int, _c_int, _c_void_p, _c_void_p],
    ),
    "mgh_dft_slice": (
        _c_int,
        [_c_int, _c_void_p, _c_ll, _c_int, _c_int, _c_int, _c_int, _c_void_p, _c_void_p, _c_int,
         _c_int, _c_void_p, _c_int, _c_int, _c_void_p, _c_void_p, _c_void_p, _c_ll, _c_int, _c_int,
         _c_void_p],
    ),
}


class HipBackendUnavailable(RuntimeError):
    """libmgh.so is missing or cannot be loaded (build it with magnelio_amd.build)."""


class MghError(RuntimeError):
    """A libmgh entry point returned a non-zero status."""


_STATUS_EXC = {1: ValueError, 2: MghError, 3: ValueError}

_lib = None


def _preload_torch_hip() -> None:
    try:
        import torch  # noqa: PLC0415
    except ImportError as exc:  # pragma: no cover - torch is part of the image
        raise HipBackendUnavailable("PyTorch-ROCm is required for the HIP backend") from exc
    rt = Path(torch.__file__).resolve().parent / "lib" / "libamdhip64.so"
    if rt.exists():
        ctypes.CDLL(str(rt), mode=ctypes.RTLD_GLOBAL)
    else:  # a system ROCm runtime
        ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)


def load() -> ctypes.CDLL:
    """Load (once) and return the library with typed signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HipBackendUnavailable(
            f"{LIB_PATH} is not built; run `python -m magnelio_amd.build` first"
        )
    _preload_torch_hip()
    try:
        lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    except OSError as exc:
        raise HipBackendUnavailable(f"cannot load {LIB_PATH}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    msg = load().mgh_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(status: int, what: str) -> None:
    """Raise the mapped Python exception for a non-zero libmgh status."""
    if status != 0:
        exc = _STATUS_EXC.get(status, MghError)
        raise exc(f"{what}: {last_error()} (status {status})")


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise on failure."""
    if name not in SIGNATURES:
        # without argtypes ctypes would pass 64-bit pointers as C ints
        raise MghError(f"{name} has no declared ctypes signature")
    fn = getattr(load(), name)
    check(fn(*args), name)
