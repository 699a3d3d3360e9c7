"""``backend="hip"``: the counterpart of the reference's backend resolver.

Reference: ``resolve_backend`` (src/magnelio/_backend/array_api.py:68-134)
maps "numpy" / "cupy" / "auto" to an array module, "auto" probing for CuPy
once per process and falling back to NumPy silently.  This package serves
one backend, "hip", and has no CPU path: "hip" returns this package after
checking that libmgh.so loads and a ROCm device is visible, and "auto"
means "hip" here — raising, never falling back, when it is not usable (the
caller's own resolver keeps its NumPy tier).  ``MAGNELIO_BACKEND``, when
set, must name "hip" or "auto".
"""

from __future__ import annotations

import os

VALID = ("hip", "auto")
_PROBED: dict = {}


class HipDeviceUnavailable(RuntimeError):
    """No usable ROCm device for backend="hip"."""


def _probe():
    from . import _lib  # noqa: PLC0415

    _lib.load()  # raises HipBackendUnavailable when the library is missing
    import torch  # noqa: PLC0415

    if not torch.cuda.is_available() or torch.cuda.device_count() < 1:
        raise HipDeviceUnavailable("backend='hip' requested but no ROCm device is visible")
    torch.zeros(1, device="cuda:0")  # context creation: the real probe


def resolve_backend(backend: str = "auto"):
    """Return this package for "hip" (or "auto") when the HIP engine can
    run; raise ``HipBackendUnavailable`` / ``HipDeviceUnavailable``
    otherwise, and ``ValueError`` for backends this package does not serve."""
    name = str(backend).strip().lower()
    if name == "auto":
        env = os.environ.get("MAGNELIO_BACKEND", "").strip().lower()
        if env and env not in VALID:
            raise ValueError(f"MAGNELIO_BACKEND={env!r} is not served by magnelio_amd; use 'hip' or 'auto'")
        name = "hip"
    if name != "hip":
        raise ValueError(f"Unknown backend: {backend!r}. magnelio_amd serves {VALID}")
    if "error" in _PROBED:
        raise _PROBED["error"]
    if not _PROBED:
        try:
            _probe()
        except Exception as exc:
            _PROBED["error"] = exc
            raise
        _PROBED["ok"] = True
    import magnelio_amd  # noqa: PLC0415

    return magnelio_amd


def is_gpu_backend(xp) -> bool:
    """The reference's GPU test (``xp is not np``, fit_td.py:312) for this backend."""
    import numpy as np  # noqa: PLC0415

    return xp is not np
